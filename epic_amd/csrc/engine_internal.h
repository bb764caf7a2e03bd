// engine_internal.h -- entry points of engine.cpp shared with the drop-in
// (dropin.cpp) and the allocation-free step (noalloc.cpp); not part of the
// exported C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/epic_amd.h"
#include "dev_common.h"
#include "kernels.h"
#include "tables.h"

int engine_add_member(epic_engine *E, const epic_case *c, const double *hdry_src, bool prime);
int engine_sync_member(epic_engine *E, int slot);
int engine_richardson(epic_engine *E, int slot);
int engine_check_errors(epic_engine *E);
const epic::MemberDesc *engine_member_desc(epic_engine *E, int slot);
hipStream_t engine_stream(epic_engine *E);
char *engine_errbuf(epic_engine *E);

// One array of a member's descriptor: the descriptor's pointer to set, its
// size, its initial contents (null: zero) and, for an array that callers of
// the engine read or write, its field id (EPIC_F_*, the tendency planes as
// EPIC_F_DUDT0 / EPIC_F_DVDT0 with three planes each), else -1.
struct MemberPart {
  void **dst;
  size_t bytes;
  const void *src;
  int field;
};
// Host tables a member's parts point into until they are copied.
struct MemberHold {
  std::vector<double> cl_h, cl4_h, clon_h, t2_h, minv_v, ninv_h;
  std::vector<int> t2_kay, flag_h, flag_v;
  std::vector<float> twf32, twi32, trf32, tri32;
};
// A second stream (default priority) and the two events that fork a branch
// of the step onto it and join it back.
struct SideStream {
  hipStream_t s2;
  hipEvent_t fork, join;
};
// 0 when the engine implements the case, else -1 with the reason in err.
int member_check(const epic_case *c, int precision, char *err, size_t cap);
// The descriptor's scalars and the list of its arrays, in the blob's order.
void member_plan(const epic_case *c, bool mixed, const epic_tables &T, epic::MemberDesc &d,
                 std::vector<MemberPart> &parts, MemberHold &hold);
// Queue one bucket's SYNC_DIAGS_ONLY sequence (sync_only) or one step on s;
// hv, divd, adjust: some member has del^4 hyperviscosity, divergence
// damping, a layer-mean heating removal.
void member_record(hipStream_t s, const epic::Bucket &b, bool sync_only, bool hv, bool divd, bool adjust);
