// Environment map light: host-side helpers shared by the scene and the pass.
#pragma once
#include "scene_host.hpp"

struct plt_scene;

namespace plt {

// The importance map of an environment map (its mip chain), from its image.
void build_env_importance(HostEnvMap& e);

}  // namespace plt
