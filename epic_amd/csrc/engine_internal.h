// engine_internal.h -- entry points of engine.cpp shared with the drop-in
// (dropin.cpp); not part of the exported C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/epic_amd.h"
#include "dev_common.h"

int engine_add_member(epic_engine *E, const epic_case *c, const double *hdry_src, bool prime);
int engine_sync_member(epic_engine *E, int slot);
int engine_richardson(epic_engine *E, int slot);
int engine_check_errors(epic_engine *E);
const epic::MemberDesc *engine_member_desc(epic_engine *E, int slot);
hipStream_t engine_stream(epic_engine *E);
char *engine_errbuf(epic_engine *E);
