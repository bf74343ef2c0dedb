// md.hip -- device-resident velocity Verlet and order-restoring copies.
//
// The reference leaves molecular dynamics to ASE (doc/running.rst:88); the
// integrator here follows ASE's VelocityVerlet step: p += dt/2 F,
// r += dt p/m, F = F(r), p += dt/2 F (the second half-kick is fused into the
// force gather kernel).
#include "system.h"

namespace {

__global__ void k_drift(int n, double dt, double half_dt, const double4 *__restrict__ frc,
                        double4 *__restrict__ mom, double4 *__restrict__ pos)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 f = frc[i];
    double4 p = mom[i];
    p.x += half_dt * f.x;
    p.y += half_dt * f.y;
    p.z += half_dt * f.z;
    mom[i] = p;
    double4 r = pos[i];
    double s = dt / p.w;  // p.w carries the mass
    r.x += s * p.x;
    r.y += s * p.y;
    r.z += s * p.z;
    pos[i] = r;
}

__global__ void k_scatter_orig(int n, const int *__restrict__ id, const double4 *__restrict__ src,
                               double4 *__restrict__ dst)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    dst[id[s]] = src[s];
}

}  // namespace

void md_drift(System &S, double dt, double half_dt)
{
    hipEvent_t ev;
    S.timer.begin("vv_drift", S.stream, &ev);
    k_drift<<<div_up(S.n, 256), 256, 0, S.stream>>>(S.n, dt, half_dt, S.frc.p, S.mom.p, S.pos.p);
    HIP_TRY(hipGetLastError());
    S.timer.end("vv_drift", S.stream, ev);
}

void sys_scatter_original(System &S, const double4 *src, double4 *dst)
{
    k_scatter_orig<<<div_up(S.n, 256), 256, 0, S.stream>>>(S.n, S.id.p, src, dst);
    HIP_TRY(hipGetLastError());
}
