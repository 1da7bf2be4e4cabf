#include "bgpu_internal.h"
namespace bgpu {
void destroy_scheduler(Scheduler*) {}
int build_pt_data(Ctx*, const bgpu_scene_desc* d, Scene* sc) {
  sc->pt.shape_light_host.assign(d->num_shapes, -1);
  return BGPU_OK;
}
}  // namespace bgpu
extern "C" {
int bgpu_render(bgpu_scene*, const bgpu_job_desc*, float*, float*) { return bgpu::fail(BGPU_E_STATE, "not built"); }
int bgpu_job_submit(bgpu_scene*, const bgpu_job_desc*, bgpu_job**) { return bgpu::fail(BGPU_E_STATE, "not built"); }
int bgpu_job_poll(bgpu_job*, float*, float*, int32_t*, int32_t*) { return bgpu::fail(BGPU_E_STATE, "not built"); }
int bgpu_job_wait(bgpu_job*, double) { return bgpu::fail(BGPU_E_STATE, "not built"); }
int bgpu_job_cancel(bgpu_job*) { return bgpu::fail(BGPU_E_STATE, "not built"); }
int bgpu_job_release(bgpu_job*) { return bgpu::fail(BGPU_E_STATE, "not built"); }
int bgpu_job_latency_ms(bgpu_job*, double*) { return bgpu::fail(BGPU_E_STATE, "not built"); }
}
