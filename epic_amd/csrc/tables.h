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
} epic_tables;
int  epic_tables_build(const epic_case *c, epic_tables *T);
void epic_tables_free(epic_tables *T);
#ifdef __cplusplus
}
#endif
#endif
