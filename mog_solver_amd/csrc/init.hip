// Initial conditions: Perlin-fBm rejection sampler on the GPU.
//
// Same algorithm as shaders/init_particles.glsl:20-287 (3D branch), written
// for HIP: a 32-bit integer hash drives per-particle candidates, a 4-octave
// fractal Perlin field gives the acceptance density
// pow(clamp(fbm + 0.5, 0, 1), contrast), up to 64 attempts per particle with
// a uniform fallback; velocities are uniform in [-W/2, W/2) per axis and the
// mass is base_mass.  Integer steps are exact; float steps are float32 with
// the device's transcendental functions, so a candidate whose density lies
// within a few ulps of its threshold may be decided differently than by
// another implementation.
#include "mog_common.h"

namespace mog {

__device__ __forceinline__ unsigned ihash(unsigned x) {
  x ^= x >> 16;
  x *= 0x85ebca6bu;
  x ^= x >> 13;
  x *= 0xc2b2ae35u;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ unsigned ihash2(unsigned x, unsigned y) {
  return ihash(x ^ (ihash(y) + 0x9e3779b9u + (x << 6) + (x >> 2)));
}
__device__ __forceinline__ unsigned ihash3(unsigned x, unsigned y, unsigned z) {
  return ihash(ihash2(x, y) ^ (ihash(z) + 0x9e3779b9u));
}
// float(h) / 4294967295.0 with float32 operands (the constant rounds to 2^32)
__device__ __forceinline__ float urand(unsigned s) { return (float)ihash(s) / 4294967296.0f; }

__device__ __forceinline__ float3 urand3(unsigned s) {
  const unsigned h1 = ihash(s), h2 = ihash(h1);
  return make_float3(urand(s), urand(h1), urand(h2));
}

__device__ __forceinline__ float3 gradient(int x, int y, int z, unsigned s) {
  const unsigned h = ihash3((unsigned)x + s, (unsigned)y + s * 7u, (unsigned)z + s * 13u);
  const float theta = ((float)(h & 0xFFFFu) * 6.28318530718f) / 65535.0f;
  const float phi = acosf(((2.0f * (float)(h >> 16)) / 65535.0f) - 1.0f);
  const float sp = sinf(phi);
  return make_float3(sp * cosf(theta), sp * sinf(theta), cosf(phi));
}

__device__ __forceinline__ float fade(float t) {
  return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
}
// GLSL mix(x, y, a) = x * (1 - a) + y * a
__device__ __forceinline__ float lerp_glsl(float x, float y, float a) {
  return x * (1.0f - a) + y * a;
}

__device__ float perlin(float px, float py, float pz, unsigned s) {
  const float fx0 = floorf(px), fy0 = floorf(py), fz0 = floorf(pz);
  const int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
  const float fx = px - fx0, fy = py - fy0, fz = pz - fz0;
  const float ux = fade(fx), uy = fade(fy), uz = fade(fz);
  float n[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
    const float3 gr = gradient(ix + dx, iy + dy, iz + dz, s);
    const float ox = fx - (float)dx, oy = fy - (float)dy, oz = fz - (float)dz;
    n[c] = gr.x * ox + gr.y * oy + gr.z * oz;
  }
  const float nx00 = lerp_glsl(n[0], n[1], ux), nx10 = lerp_glsl(n[2], n[3], ux);
  const float nx01 = lerp_glsl(n[4], n[5], ux), nx11 = lerp_glsl(n[6], n[7], ux);
  const float nxy0 = lerp_glsl(nx00, nx10, uy), nxy1 = lerp_glsl(nx01, nx11, uy);
  return lerp_glsl(nxy0, nxy1, uz);
}

__device__ float fbm(float px, float py, float pz, unsigned s) {
  float value = 0.0f, amplitude = 0.5f, frequency = 1.0f;
  for (int i = 0; i < 4; ++i) {
    value += amplitude * perlin(px * frequency, py * frequency, pz * frequency, s + (unsigned)i * 1000u);
    amplitude *= 0.5f;
    frequency *= 2.0f;
  }
  return value;
}

struct InitParams {
  float W, noise_scale, contrast, base_mass, spawn_buffer;
  unsigned seed;
  int sphere;
};

__global__ __launch_bounds__(256) void init_perlin_kernel(float4* __restrict__ pos,
                                                          float4* __restrict__ vel, long long n,
                                                          InitParams q) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  const unsigned ps = q.seed + (unsigned)gid * 12345u;
  const float minb = q.spawn_buffer * q.W;
  const float maxb = (1.0f - q.spawn_buffer) * q.W;
  const float range = maxb - minb;
  const float center = q.W * 0.5f;
  const float radius = range * 0.5f;
  float3 p = make_float3(0.f, 0.f, 0.f);
  unsigned attempts = 0;
  for (; attempts < 64u; ++attempts) {
    float3 cand;
    if (q.sphere) {
      float3 sp;
      unsigned sa = 0;
      do {
        const float3 r = urand3(ps + attempts * 1000u + sa * 100u);
        sp = make_float3(r.x * 2.0f - 1.0f, r.y * 2.0f - 1.0f, r.z * 2.0f - 1.0f);
        ++sa;
      } while ((sp.x * sp.x + sp.y * sp.y + sp.z * sp.z) > 1.0f && sa < 16u);
      cand = make_float3(center + sp.x * radius, center + sp.y * radius, center + sp.z * radius);
    } else {
      const float3 r = urand3(ps + attempts * 1000u);
      cand = make_float3(r.x * range + minb, r.y * range + minb, r.z * range + minb);
    }
    float d = fbm((cand.x * q.noise_scale) / q.W, (cand.y * q.noise_scale) / q.W,
                  (cand.z * q.noise_scale) / q.W, q.seed);
    d = d + 0.5f;
    d = powf(fminf(fmaxf(d, 0.0f), 1.0f), q.contrast);
    if (urand(ps + attempts * 1000u + 500u) < d) {
      p = cand;
      break;
    }
  }
  if (attempts >= 64u) {
    if (q.sphere) {
      const float r = powf(urand(ps), 1.0f / 3.0f) * radius;
      const float theta = urand(ihash(ps)) * 6.28318530718f;
      const float phi = acosf(2.0f * urand(ihash(ihash(ps))) - 1.0f);
      p = make_float3(center + sinf(phi) * cosf(theta) * r, center + sinf(phi) * sinf(theta) * r,
                      center + cosf(phi) * r);
    } else {
      const float3 r = urand3(ps);
      p = make_float3(r.x * range + minb, r.y * range + minb, r.z * range + minb);
    }
  }
  const float3 v = urand3(ps + 999u);
  pos[gid] = make_float4(p.x, p.y, p.z, q.base_mass);
  vel[gid] = make_float4((v.x - 0.5f) * 1.0f * q.W, (v.y - 0.5f) * 1.0f * q.W,
                         (v.z - 0.5f) * 1.0f * q.W, 0.0f);
}

hipError_t launch_init_perlin(float4* pos, float4* vel, long long n, float W, float noise_scale,
                              float contrast, unsigned seed, float base_mass, float spawn_buffer,
                              int sphere, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  InitParams q{W, noise_scale, contrast, base_mass, spawn_buffer, seed, sphere};
  hipLaunchKernelGGL(init_perlin_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos,
                     vel, n, q);
  return hipGetLastError();
}

}  // namespace mog
