// Everything generated material code may reference: node library, shading
// context, lights.  The generated text is compiled after this header by
// hiprtc (GPU module) and by g++ (CPU oracle).
#pragma once
#include "common.h"
#include "elem.h"
#include "rng.h"
#include "warp.h"
#include "bsdf.h"
#include "scene.h"
#include "shading.h"
#include "noise.h"
#include "nodes.h"
#include "lights.h"

// out-of-line material functions (large graph sets): bounds the shading
// kernel's code size and compile time
#if defined(RM_GPU)
#define RM_MATERIAL_FN_NOINLINE __host__ __device__ __noinline__
#else
#define RM_MATERIAL_FN_NOINLINE __attribute__((noinline))
#endif
