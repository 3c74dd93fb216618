/* Test driver: offsets and sizes of the structures of include/lblx.h as
 * the C compiler lays them out, in the format of abi_layout.f90. */
#include <stddef.h>
#include <stdio.h>

#include "../../include/lblx.h"

#define F(T, M) printf("off %s %s %zu %zu\n", #T, #M, offsetof(T, M), sizeof(((T *)0)->M))

int main(void)
{
  F(lblx_run, v1); F(lblx_run, v2); F(lblx_run, sample); F(lblx_run, alfal0); F(lblx_run, avmass);
  F(lblx_run, dptmin); F(lblx_run, dptfac); F(lblx_run, ilblf4); F(lblx_run, icntnm); F(lblx_run, irad);
  F(lblx_run, precision); F(lblx_run, cntscl); F(lblx_run, ifrgnx); F(lblx_run, iaersl); F(lblx_run, voigt);
  F(lblx_run, ibrd);
  printf("size lblx_run %zu\n", sizeof(lblx_run));
  F(lblx_layer, pave); F(lblx_layer, tave); F(lblx_layer, wk); F(lblx_layer, wbroad); F(lblx_layer, dv);
  F(lblx_layer, v1); F(lblx_layer, v2); F(lblx_layer, nmol); F(lblx_layer, jrad);
  printf("size lblx_layer %zu\n", sizeof(lblx_layer));
  F(lblx_stats, nlin); F(lblx_stats, npts); F(lblx_stats, nwin4); F(lblx_stats, nfix); F(lblx_stats, fixseq);
  printf("size lblx_stats %zu\n", sizeof(lblx_stats));
  F(lblx_scan, hwhm); F(lblx_scan, param); F(lblx_scan, jemit); F(lblx_scan, nofix);
  printf("size lblx_scan %zu\n", sizeof(lblx_scan));
  F(lblx_scanfn, hwhm); F(lblx_scanfn, sampl); F(lblx_scanfn, jemit); F(lblx_scanfn, npanel);
  printf("size lblx_scanfn %zu\n", sizeof(lblx_scanfn));
  return 0;
}
