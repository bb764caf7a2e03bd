/* case_internal.h -- the parts of the case builder (epic_case.c) that the
 * epic.nc reader (epic_nc.c) shares. */
#ifndef EPIC_CASE_INTERNAL_H
#define EPIC_CASE_INTERNAL_H
#include "../../include/epic_amd.h"
#ifdef __cplusplus
extern "C" {
#endif
/* the reference's advection-scheme names (src/shared/epic_initial.c:146-149),
 * indexed by epic_case_spec.adv_mass / adv_nonmass */
#define EPIC_ADV_HSU "Predictor-corrector (Hsu, Konor, and Arakawa)"
#define EPIC_ADV_AKIMA "Monotonized, 4th-order centered (Akima)"
void epic_case_alloc(epic_case *c);
void epic_case_metrics(epic_case *c);
int  epic_case_dsgth(epic_case *c);
#ifdef __cplusplus
}
#endif
#endif
