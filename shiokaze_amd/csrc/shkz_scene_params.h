// shkz_scene_params.h -- the shapes of a synthetic scene as plain data, shared by the host
// generator (shkz_scenes.cpp) and the device generator (shkz_synth.hip).  Internal to the library.
#pragma once
#include <stdint.h>

#include "../../include/shkz_mi355x.h"

struct ShkzSceneParams {
    int32_t n_liquid, n_solid, has_box, all_fluid;  // all_fluid: phi = -1 everywhere, no push-out
    double pool;            // liquid below y < pool; < 0 = no pool
    double liquid[8][4];    // sphere centre xyz, radius
    double solid[4][4];
    double box0[3], box1[3];
    double shift[3];        // rigid translation of the liquid level set
};

#ifdef __cplusplus
extern "C"
#endif
    __attribute__((visibility("hidden"))) int
    shkz_scene_params_internal(const shkz_scene_desc *desc, ShkzSceneParams *out);
