/* ref_pv_al81.c -- test infrastructure: compiles one of the reference's
 * sources (REF_SOURCE, named by the Makefile, read where it lies under the
 * reference tree) with PV_SCHEME set to ARAKAWA_LAMB_1981, the other value
 * that include/epic.h:79-86 offers for it ("Define PV_SCHEME to be
 * SADOURNEY_1975 or ARAKAWA_LAMB_1981").  epic.h fixes the macro with no
 * guard, so the value is set between its inclusion and the source's; the
 * source's own #include <epic.h> is then a no-op and its
 * #include <epic_pv_schemes.h> takes the Arakawa-Lamb weights.  Only the
 * sources that read PV_SCHEME in the step are built this way
 * (epic_timestep.c: uv_core; epic_funcs_diag.c: get_kin, gz_from_u). */
#include <epic.h>
#undef PV_SCHEME
#define PV_SCHEME ARAKAWA_LAMB_1981
#include REF_SOURCE
