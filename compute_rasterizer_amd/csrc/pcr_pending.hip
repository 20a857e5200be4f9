// pcr_pending.hip — entry points whose implementation is not built yet; each
// returns PCR_E_INVALID with an explanatory error.
#include "pcr_internal.h"
using namespace pcr;
extern "C" {
#define PENDING(name) set_error(#name ": not implemented in this build"); return PCR_E_INVALID;
int pcr_cloud_build_octree_host(pcr_ctx*, const float*, const uint32_t*, uint64_t, const pcr_octree_opts*, pcr_cloud**) { PENDING(pcr_cloud_build_octree_host) }
int pcr_cloud_build_octree_device(pcr_ctx*, const float*, const uint32_t*, uint64_t, const pcr_octree_opts*, pcr_cloud**) { PENDING(pcr_cloud_build_octree_device) }
int pcr_lod_selection(pcr_ctx*, pcr_cloud*, uint8_t*, uint32_t) { PENDING(pcr_lod_selection) }
int pcr_multiview_create(pcr_ctx*, pcr_cloud*, const pcr_view*, int, pcr_fb* const*, const pcr_render_params*, const pcr_tune*, int, pcr_multiview**) { PENDING(pcr_multiview_create) }
int pcr_multiview_set_views(pcr_multiview*, const pcr_view*, int) { PENDING(pcr_multiview_set_views) }
int pcr_multiview_launch(pcr_multiview*) { PENDING(pcr_multiview_launch) }
int pcr_multiview_destroy(pcr_multiview*) { PENDING(pcr_multiview_destroy) }
int pcr_autotune(pcr_ctx*, pcr_cloud*, const pcr_view*, int, pcr_fb* const*, const pcr_render_params*, int, pcr_tune*, float*) { PENDING(pcr_autotune) }
int pcr_server_create(pcr_ctx*, uint64_t, int, int, pcr_server**) { PENDING(pcr_server_create) }
int pcr_server_add_cloud(pcr_server*, pcr_cloud*, int*) { PENDING(pcr_server_add_cloud) }
int pcr_server_submit(pcr_server*, const pcr_request*, uint64_t*) { PENDING(pcr_server_submit) }
int pcr_server_poll(pcr_server*, uint64_t, uint32_t*, double*) { PENDING(pcr_server_poll) }
int pcr_server_wait(pcr_server*, uint64_t, uint32_t*, double*) { PENDING(pcr_server_wait) }
int pcr_server_cancel(pcr_server*, uint64_t) { PENDING(pcr_server_cancel) }
int pcr_server_stats_get(pcr_server*, pcr_server_stats*) { PENDING(pcr_server_stats_get) }
int pcr_server_destroy(pcr_server*) { PENDING(pcr_server_destroy) }
}
