// Integrator hooks; the wavefront integrator lives in render.hip.
#include "engine_internal.h"

vt_status engine_on_scene_changed(vt_engine* e) {
  e->samples_accumulated = 0;
  return VT_OK;
}

void engine_release_render(vt_engine*) {}
