/* tables.h -- run-once host tables of the engine (see epic_tables.c). */
#ifndef EPIC_TABLES_H
#define EPIC_TABLES_H
#include "../../include/epic_amd.h"
#ifdef __cplusplus
extern "C" {
#endif
typedef struct {
  double *mn_inv, *hl_h, *hl_pv, *hc[3], *vc[3], *twf, *twi, *trf, *tri;
  int *rows_h, *rows_v;
  int n_mn_inv, n_hl, n_tw, n_tr, nrow_h, nrow_v;
  double coef2, coef4, coef6, coef8, m0z, nudiv, eps13, gcoef, mu3, mu2;
  /* Newtonian cooling toward a sounding (epic_case_table_cooling):
   * snd = {-log p, T, dT/d(-log p)} of the sounding, ironed; cool = {-log p,
   * cooling time, its derivative}; n_snd = n_cool = 0 without it */
  double *snd[3], *cool[3];
  int n_snd, n_cool;
} epic_tables;
/* whether the case relaxes toward its sounding with the tabulated cooling
 * time (newt_cool_on on any planet but Held-Suarez, with a sounding) */
int  epic_case_table_cooling(const epic_case *c);
int  epic_tables_build(const epic_case *c, epic_tables *T);
void epic_tables_free(epic_tables *T);
#ifdef __cplusplus
}
#endif
#endif
