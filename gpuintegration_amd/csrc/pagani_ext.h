// pagani_ext.h -- library-internal entry points shared between the driver
// and the precompiled Genz evaluation (genz_eval.hip).
#pragma once

#include "pagani.h"

#ifdef __cplusplus
extern "C" {
#endif

// Launch the evaluation of a Genz descriptor on args->stream.
int pagani_genz_eval_launch(const pagani_genz* f, int precision, const pagani_eval_args* args);

#ifdef __cplusplus
}
#endif
