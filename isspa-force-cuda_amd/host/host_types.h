// Plain vector types and constants of the host layer.  The reference's host
// classes use CUDA's float4/float2/int2/int4; this layer is compiled with a
// host C++ compiler against the C ABI only, so it declares layout-identical
// PODs of its own (16/8/8/16 bytes, the device layouts of include/isspa_amd.h).
#pragma once

#include <cstdint>

struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
struct int2 { int x, y; };
struct int4 { int x, y, z, w; };

static_assert(sizeof(float4) == 16 && sizeof(float2) == 8 && sizeof(int4) == 16 &&
                  sizeof(int2) == 8,
              "device layouts");

// src/constants.h:2-5
#define MAXCHAR 500
#define LONGCHAR 2400
#define MAXLEN 80
#define nDim 3
