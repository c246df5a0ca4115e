/*
 * force.hip -- per-particle force factors and the SPH pressure + artificial viscosity pair
 * sums over the symmetric neighbour-leaf lists.
 *
 * Restates src/force.cu:92-137 (forcePrepKernel) and :145-383 (sphForceKernelJList) for
 * gfx950: one wavefront per i-leaf over its symmetric list.  Only x, y, z and h of the
 * j-particles go through the LDS tile, because the cheap reject (r^2 >= R^2 max(hi^2,hj^2))
 * removes most candidates (~95% in the reference's measurements, force.cu:242-244); the
 * remaining j-fields are read only for a pair inside either support.
 */
#include "density.hpp"
#include "tile.hpp"

namespace csph
{

constexpr int kForceWaves   = 4;
constexpr int kForceThreads = kForceWaves * kWave;

__global__ __launch_bounds__(256) void forcePrepKernel(const double* __restrict__ h, const double* __restrict__ gradh,
                                                       const double* __restrict__ pro2, double* __restrict__ hsqinv,
                                                       double* __restrict__ hinv, double* __restrict__ rhoh,
                                                       double* __restrict__ rho1, double* __restrict__ grkfac,
                                                       double* __restrict__ pres, double* __restrict__ auterm,
                                                       double* __restrict__ divfac, int n, double pmass, double alphau,
                                                       double hfact)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double hi = h[i];
    if (!(hi > 0.0))
    {
        hsqinv[i] = 0.0; hinv[i] = 0.0; rhoh[i] = 0.0; rho1[i] = 0.0; grkfac[i] = 0.0;
        pres[i] = 0.0; auterm[i] = 0.0; divfac[i] = 0.0;
        return;
    }
    const double hs = 1.0 / (hi * hi), h4 = hs * hs, hfoh = hfact / hi;
    const double rho = pmass * hfoh * hfoh * hfoh;
    const double dhdrho = -hi / (3.0 * rho);
    const double omega = 1.0 - dhdrho * gradh[i];
    const double oinv = (omega > 0.0) ? (1 / omega) : 1.0;
    hsqinv[i] = hs;
    hinv[i]   = 1.0 / hi;
    rhoh[i]   = rho;
    rho1[i]   = 1.0 / rho;
    grkfac[i] = h4 * kCnormk * oinv;
    pres[i]   = pro2[i] * rho * rho;
    auterm[i] = 0.5 * pmass * (1.0 / rho) * alphau;
    divfac[i] = kCnormk * oinv * h4 / rho;
}

// One wavefront per i-leaf, the leaf's particles up to 64 at a time, g = floor(64 / c) lanes
// per particle splitting every tile (lane s of a group takes slots s, s + g, ...), then a group
// reduction; small leaves keep the wavefront busy.  A list entry is staged only if its leaf box
// lies within R max(h_i, hmax of the j-leaf) of some particle of the chunk: a dropped leaf holds
// no pair inside either support, so the pair sums are unchanged (for g = 1 even their order).
template<bool PER, bool DISC, bool FULL>
__global__ __launch_bounds__(kForceThreads) void forceKernel(ForceArgs A)
{
    __shared__ int    segStart[kForceWaves][kWave];
    __shared__ int    segJ[kForceWaves][kWave];
    __shared__ double tx[kForceWaves][kWave], ty[kForceWaves][kWave], tz[kForceWaves][kWave];
    __shared__ double th[kForceWaves][kWave];
    __shared__ int    tj[kForceWaves][kWave];
    __shared__ double cp[kForceWaves][4][kWave];   // the chunk's particles: x, y, z, R h
    const int w = threadIdx.x >> 6, lane = laneId();
    const int a = blockIdx.x * kForceWaves + w;
    unsigned long long nAcc = 0, nCand = 0;
    if (a < A.nLeaves)
    {
        const int iLeaf = a;
        const uint32_t p0 = A.layout[iLeaf], p1 = A.layout[iLeaf + 1];
        const uint32_t lo = A.jOffset[a], hi = A.jOffset[a + 1];
        const DBox box = A.boxes[A.leafSet[iLeaf]];
        const double pmass = A.pmass, beta = A.beta;
        for (uint32_t q0 = p0; q0 < p1; q0 += kWave)
        {
            const int c = __builtin_amdgcn_readfirstlane((int)min((uint32_t)kWave, p1 - q0));
            const int g = __builtin_amdgcn_readfirstlane(kWave / c);
            const int grp = lane / g, sub = lane - grp * g;
            const bool has = grp < c;
            const uint32_t i = q0 + (uint32_t)(has ? grp : 0);
            const double xi = A.x[i], yi = A.y[i], zi = A.z[i], hi_ = A.h[i];
            const double vxi = A.vx[i], vyi = A.vy[i], vzi = A.vz[i];
            const double hsi = A.hsqinv[i], rhoi = A.rhoh[i], gki = A.grkfac[i], hinvi = A.hinv[i];
            const double pro2i = A.pro2[i], vwi = A.cs[i], ali = A.alpha[i], eni = A.u[i];
            const double rho1i = A.rho1[i], pri = A.pres[i], aui = A.auterm[i];
            if (lane < c)
            {
                const uint32_t ic = q0 + (uint32_t)lane;
                cp[w][0][lane] = A.x[ic]; cp[w][1][lane] = A.y[ic]; cp[w][2][lane] = A.z[ic];
                cp[w][3][lane] = kRad * A.h[ic];
            }
            waveLdsSync();
            double ix = 0, iy = 0, iz = 0, jx = 0, jy = 0, jz = 0, f4s = 0, vmax = 0, divs = 0;
            int acc = 0;
            uint32_t cand = 0;
            auto keep = [&](int jl) {
                const int node = A.leafNode[jl];
                double hj = A.hmaxLeaf[jl];
                const double rj = (isfinite(hj) && hj > 0.0) ? kRad * hj : 0.0;
                for (int k = 0; k < c; ++k)
                {
                    const double r = fmax(cp[w][3][k], rj);
                    if (pointNodeGap2<PER>(cp[w][0][k], cp[w][1][k], cp[w][2][k], A.cx, A.cy, A.cz, A.sx, A.sy, A.sz,
                                           node, box) <= r * r * kCullSlack)
                        return true;
                }
                return false;
            };
            auto load = [&](int l, int j) {
                tx[w][l] = A.x[j]; ty[w][l] = A.y[j]; tz[w][l] = A.z[j]; th[w][l] = A.h[j]; tj[w][l] = j;
            };
            auto pair = [&](int k) {
                double dx = xi - tx[w][k], dy = yi - ty[w][k], dz = zi - tz[w][k];
                if constexpr (PER)
                {
                    if (fabs(dx) > 0.5 * box.l[0]) dx = dx - box.l[0] * copysign(1.0, dx);
                    if (fabs(dy) > 0.5 * box.l[1]) dy = dy - box.l[1] * copysign(1.0, dy);
                    if (fabs(dz) > 0.5 * box.l[2]) dz = dz - box.l[2] * copysign(1.0, dz);
                }
                const double dr2 = dx * dx + dy * dy + dz * dz;
                const double hj = th[w][k];
                const double rmax2 = kRad2 * fmax(hi_ * hi_, hj * hj) * (1.0 + 1e-10);
                if (!((dr2 > 0.0) && (dr2 < rmax2))) return;
                const int j = tj[w][k];
                const double qi2 = dr2 * hsi, qj2 = dr2 * A.hsqinv[j];
                const double dr = sqrt(dr2), rij1 = 1.0 / dr;
                const double rx = dx * rij1, ry = dy * rij1, rz = dz * rij1;
                const double dvx = vxi - A.vx[j], dvy = vyi - A.vy[j], dvz = vzi - A.vz[j];
                const double projv = dvx * rx + dvy * ry + dvz * rz;
                const double rhoj = A.rhoh[j], rho1j = A.rho1[j], pro2j = A.pro2[j];
                const double vwj = A.cs[j], alj = A.alpha[j], enj = A.u[j], prj = A.pres[j], auj = A.auterm[j];
                const double vsigi = fmax(vwi - beta * projv, 0.0);
                const double vsigavi = fmax(ali * vwi - beta * projv, 0.0);
                const double vsigj = fmax(vwj - beta * projv, 0.0);
                const double vsigavj = fmax(alj * vwj - beta * projv, 0.0);
                const double pvmax = fmax(vsigi, vsigj);
                double qr2i = 0.0, qr2j = 0.0;
                const double denij = eni - enj;
                const double vsigu = sqrt(fabs(pri - prj) * (2.0 / (rhoi + rhoj)));
                if constexpr (DISC)
                {
                    const double hori = hi_ * rij1, horj = hj * rij1;
                    if (projv < 0.0)
                    {
                        qr2i = -0.5 * rho1i * (ali * vwi - beta * projv) * hori * projv;
                        qr2j = -0.5 * rho1j * (alj * vwj - beta * projv) * horj * projv;
                    }
                    else
                    {
                        qr2i = -0.5 * rho1i * ali * vwi * hori * projv;
                        qr2j = -0.5 * rho1j * alj * vwj * horj * projv;
                    }
                }
                else if (projv < 0.0)
                {
                    qr2i = -0.5 * rho1i * vsigavi * projv;
                    qr2j = -0.5 * rho1j * vsigavj * projv;
                }
                bool counted = false;
                if (qi2 < kRad2)
                {
                    counted = true;
                    vmax = fmax(vmax, pvmax);
                    double wk, dwk;
                    const double q = dr * hinvi;
                    kernWdW<double>(qi2, q, wk, dwk);
                    divs += pmass * dwk * projv;
                    const double gk = dwk * gki;
                    const double gp = pmass * (pro2i + qr2i) * gk;
                    ix += -gp * rx; iy += -gp * ry; iz += -gp * rz;
                    const double pdvt = pmass * pro2i * projv * gk;
                    const double diss = DISC ? -0.5 * pmass * rho1i * ali * vwi * (hi_ * rij1) * projv * projv * gk
                                             : pmass * qr2i * projv * gk;
                    const double dend = vsigu * denij * aui * gk;
                    if constexpr (FULL) { f4s += pdvt; f4s += diss; }
                    else
                    {
                        if (A.pdv) f4s += pdvt;
                        if (A.shock) f4s += diss;
                    }
                    f4s += dend;
                }
                if (qj2 < kRad2)
                {
                    counted = true;
                    vmax = fmax(vmax, pvmax);
                    double wk, dwk;
                    const double q = dr * A.hinv[j];
                    kernWdW<double>(qj2, q, wk, dwk);
                    const double gk = dwk * A.grkfac[j];
                    const double gp = pmass * (pro2j + qr2j) * gk;
                    jx -= gp * rx; jy -= gp * ry; jz -= gp * rz;
                    f4s += vsigu * denij * auj * gk;
                }
                acc += counted ? 1 : 0;
            };
            // lane sub of a group takes slots sub, sub + g, ... below nt (g = 1: the list order)
            auto consume = [&](int nt) {
                if (sub == 0) cand += (uint32_t)nt;
                for (int k = sub; k < nt; k += g) pair(k);
            };
            forEachTileIf(segStart[w], segJ[w], A.layout, A.jList, lo, hi, keep, load, [](int) {}, consume);
            // group sums (and the max of vsigmax): a tree over lanes sub .. g - 1 of each group
            for (int off = 1; off < g; off <<= 1)
            {
                auto add = [&](double& v) {
                    const double o = __shfl_down(v, off, kWave);
                    if (sub + off < g) v += o;
                };
                add(ix); add(iy); add(iz); add(jx); add(jy); add(jz); add(f4s); add(divs);
                const double om = __shfl_down(vmax, off, kWave);
                if (sub + off < g) vmax = fmax(vmax, om);
                const int oa = __shfl_down(acc, off, kWave);
                if (sub + off < g) acc += oa;
            }
            if (has && sub == 0)
            {
                A.fx[i] = ix + jx; A.fy[i] = iy + jy; A.fz[i] = iz + jz;
                A.f4[i] = f4s; A.vsig[i] = vmax;
                A.divv[i] = -divs * A.divfac[i];
                nAcc += (unsigned long long)acc;
                nCand += cand;
            }
        }
    }
    nAcc = waveReduceSum(nAcc);
    nCand = waveReduceSum(nCand);
    if (laneId() == 0 && (nAcc | nCand))
    {
        unsigned long long* slot = A.counters + ((blockIdx.x * kForceWaves + w) % kCounterSlots) * kCounterStride;
        atomicAdd(slot + 0, nAcc);
        atomicAdd(slot + 1, nCand);
    }
}

__global__ void deadForceKernel(int first, int n, double* __restrict__ fx, double* __restrict__ fy,
                                double* __restrict__ fz, double* __restrict__ f4, double* __restrict__ vsig,
                                double* __restrict__ divv)
{
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fx[i] = 0.0; fy[i] = 0.0; fz[i] = 0.0; f4[i] = 0.0; vsig[i] = 0.0; divv[i] = 0.0;
}

void launchForcePrep(const double* h, const double* gradh, const double* pro2, double* hsqinv, double* hinv,
                     double* rhoh, double* rho1, double* grkfac, double* pres, double* auterm, double* divfac, int n,
                     double pmass, double alphau, double hfact, hipStream_t st)
{
    if (n <= 0) return;
    forcePrepKernel<<<iceil(n, 256), 256, 0, st>>>(h, gradh, pro2, hsqinv, hinv, rhoh, rho1, grkfac, pres, auterm,
                                                   divfac, n, pmass, alphau, hfact);
    LAUNCH_CHECK();
}

void launchForce(const ForceArgs& a, bool periodic, bool disc, bool full, hipStream_t st)
{
    if (a.nLeaves <= 0) return;
    const int grid = iceil(a.nLeaves, kForceWaves);
#define CSPH_F(P, D, F) forceKernel<P, D, F><<<grid, kForceThreads, 0, st>>>(a)
    if (periodic)
    {
        if (disc) { if (full) CSPH_F(true, true, true); else CSPH_F(true, true, false); }
        else      { if (full) CSPH_F(true, false, true); else CSPH_F(true, false, false); }
    }
    else
    {
        if (disc) { if (full) CSPH_F(false, true, true); else CSPH_F(false, true, false); }
        else      { if (full) CSPH_F(false, false, true); else CSPH_F(false, false, false); }
    }
#undef CSPH_F
    LAUNCH_CHECK();
}

void launchDeadForce(int first, int n, double* fx, double* fy, double* fz, double* f4, double* vsig, double* divv,
                     hipStream_t st)
{
    if (n > first)
    {
        deadForceKernel<<<iceil(n - first, 256), 256, 0, st>>>(first, n, fx, fy, fz, f4, vsig, divv);
        LAUNCH_CHECK();
    }
}

} // namespace csph
