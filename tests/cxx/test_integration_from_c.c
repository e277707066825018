/* C caller of the C interop shim, as test/cuda/pagani/test_integration_from_c.c
 * of the reference: integrate x*y over the unit square at epsrel 1e-6 and
 * check the result against the exact 1/4. */
#include <math.h>
#include <stdio.h>

#include "pagani.h"

int main(void)
{
  double result = 0.0;
  const int rc = do_integration_from_c(&result);
  printf("Status: %i   result: %.17g\n", rc, result);
  if (rc != 0) return rc;
  const double achieved = fabs(result - 0.25) / 0.25;
  if (achieved > 1.0e-6) {
    printf("Failed to achieve relative error of 1.0e-6 (got %g)\n", achieved);
    return 1;
  }
  return 0;
}
