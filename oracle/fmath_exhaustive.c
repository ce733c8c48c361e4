/*
 * Exhaustive check of the shared float transcendentals (test infrastructure).
 * For every float argument, compares mgx_fmath.h's expf/logf/erff/expm1f with
 * glibc's float function and with (float) of glibc's double function (the
 * correctly rounded reference), and prints the mismatch counts.
 *   make -C oracle fmath_exhaustive && oracle/fmath_exhaustive erf
 */
#include <stdio.h>
#include <math.h>
#include <string.h>
#include <stdint.h>
#include "mgx_fmath.h"
static float bf(uint32_t u){float f; memcpy(&f,&u,4); return f;}
int main(int argc, char** argv){
  const char* which = argv[1];
  long long mism_glibc=0, mism_cr=0, n=0, glibc_vs_cr=0;
  #pragma omp parallel for reduction(+:mism_glibc,mism_cr,n,glibc_vs_cr) schedule(static)
  for (long long i=0;i<(1LL<<32);i++){
    float x = bf((uint32_t)i);
    if (isnan(x)) continue;
    float a,g,c;
    if (!strcmp(which,"exp")) { if (x>89||x<-104) continue; a=mgx_expf(x); g=expf(x); c=(float)exp((double)x);}
    else if (!strcmp(which,"log")) { if (x<=0||isinf(x)) continue; a=mgx_logf(x); g=logf(x); c=(float)log((double)x);}
    else if (!strcmp(which,"erf")) { a=mgx_erff(x); g=erff(x); c=(float)erf((double)x);}
    else { if (x>89||x<-20) continue; a=mgx_expm1f(x); g=expm1f(x); c=(float)expm1((double)x);}
    n++;
    uint32_t ua,ug,uc; memcpy(&ua,&a,4); memcpy(&ug,&g,4); memcpy(&uc,&c,4);
    if (ua!=ug) mism_glibc++;
    if (ua!=uc) mism_cr++;
    if (ug!=uc) glibc_vs_cr++;
  }
  printf("%s: n=%lld  mgx!=glibc_float %lld  mgx!=(float)glibc_double %lld  glibc_float!=(float)glibc_double %lld\n", which, n, mism_glibc, mism_cr, glibc_vs_cr);
}
