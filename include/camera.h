// camera.h -- pinhole RayGenerator of the reference (camera.h:1-32 of
// Ushio/MinimumLBVH) for HIP, on the facade's vector helpers.  shoot() returns
// an unnormalised direction; lookat() takes fovy in radians.  Bit-identical to
// mlbvh_raygen_lookat / the library's render kernels when compiled with
// -ffp-contract=off.
#pragma once

#include "minimum_lbvh.h"

struct RayGenerator {
    float3 m_origin = {0.0f, 0.0f, 0.0f};
    float3 m_right = {1.0f, 0.0f, 0.0f};
    float3 m_up = {0.0f, 1.0f, 0.0f};

    __host__ __device__ void lookat(float3 eye, float3 center, float3 up, float fovy, int imageWidth,
                                    int imageHeight)
    {
        using namespace minimum_lbvh;
        float3 f = normalize(sub(center, eye));
        float3 s = normalize(cross(f, up));
        float3 u = cross(s, f);
        float tanThetaY = tanf(fovy * 0.5f);
        float tanThetaX = tanThetaY / imageHeight * imageWidth;
        m_origin = eye;
        m_right = mul(s, tanThetaX);
        m_up = mul(u, tanThetaY);
    }

    __host__ __device__ void shoot(float3* ro, float3* rd, float u, float v) const
    {
        using namespace minimum_lbvh;
        float3 from = m_origin;
        float3 forward = normalize(cross(m_up, m_right));
        float3 nr = neg(m_right), nu = neg(m_up);
        float3 lr = add(nr, mul(sub(m_right, nr), u));  // lerp(-right, right, u)
        float3 lu = add(m_up, mul(sub(nu, m_up), v));   // lerp(up, -up, v)
        float3 to = add(add(add(m_origin, forward), lr), lu);
        *ro = from;
        *rd = sub(to, from);
    }
};

static_assert(sizeof(RayGenerator) == sizeof(mlbvh_raygen), "RayGenerator is the 36-byte mlbvh_raygen");
