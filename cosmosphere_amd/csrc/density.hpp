/*
 * density.hpp -- launch interface of the density, closure, compaction and gradient kernels
 * (density.hip) and of the force kernels (force.hip).
 */
#pragma once

#include "state.hpp"

namespace csph
{

struct DensArgs
{
    const int*      leaves;     // active leaf ids (nullptr: leaf a is leaf a)
    int             nLeaves;
    const uint32_t* jOffset;    // CSR of the lists, indexed by position in `leaves`
    const int*      jList;
    const uint32_t* layout;
    const int*      leafNode;
    const int*      leafSet;
    const double*   cx; const double* cy; const double* cz;
    const double*   sx; const double* sy; const double* sz;
    const double*   hmaxLeaf;
    const double*   x; const double* y; const double* z;
    double*         h;
    double*         rho;
    double*         gradh;
    uint8_t*        active;
    const DBox*     boxes;
    double          pmass, tolh, hfact;
    unsigned long long* counters;
};

struct GradArgs
{
    int             nLeaves;
    const uint32_t* jOffset;
    const int*      jList;
    const uint32_t* layout;
    const int*      leafSet;
    const double*   x; const double* y; const double* z;
    const double*   vx; const double* vy; const double* vz;
    const double*   ax; const double* ay; const double* az;
    const double*   h;
    double*         rho; double* gradh; double* divv; double* xi; double* ddivvdt;
    const DBox*     boxes;
    double          pmass;
};

struct ForceArgs
{
    int             nLeaves;
    const uint32_t* jOffset;
    const int*      jList;
    const uint32_t* layout;
    const int*      leafSet;
    const double*   x; const double* y; const double* z;
    const double*   vx; const double* vy; const double* vz;
    const double*   h;
    const double*   pro2; const double* cs; const double* alpha; const double* u;
    const double*   hsqinv; const double* hinv; const double* rhoh; const double* rho1;
    const double*   grkfac; const double* pres; const double* auterm; const double* divfac;
    double*         fx; double* fy; double* fz; double* f4; double* vsig; double* divv;
    const DBox*     boxes;
    double          pmass, beta;
    int             pdv, shock;
    unsigned long long* counters;
};

// mode 0: NR on active particles (FLAT_PARTICLE); 1: NR on every particle of an active
// leaf (WARP_PER_LEAF); 2: closure, rho and d(rho)/d(h) at the current h, no update.
void launchDensity(const DensArgs& a, int precision, int mode, bool periodic, hipStream_t st);
void launchGradients(const GradArgs& a, bool periodic, hipStream_t st);
void launchDeadGradients(int first, int n, double* divv, double* xi, double* ddivvdt, hipStream_t st);
void launchLeafFlags(const int* leaves, int nLeaves, const uint32_t* layout, const uint8_t* active, uint32_t* flag,
                     unsigned long long* nActive, hipStream_t st);
void launchLeafScatter(const int* leaves, int nLeaves, const uint32_t* flag, const uint32_t* pos, int* out,
                       hipStream_t st);
void launchInitActive(uint8_t* active, int nAlive, int n, double* rho, double* gradh, hipStream_t st);
void launchIota(int* v, int n, hipStream_t st);
void launchMaxH(const double* h, int n, unsigned long long* out, hipStream_t st);
int  counterSlots();
int  counterStride();

void launchForcePrep(const double* h, const double* gradh, const double* pro2, double* hsqinv, double* hinv,
                     double* rhoh, double* rho1, double* grkfac, double* pres, double* auterm, double* divfac, int n,
                     double pmass, double alphau, double hfact, hipStream_t st);
void launchForce(const ForceArgs& a, bool periodic, bool disc, bool full, hipStream_t st);
void launchDeadForce(int first, int n, double* fx, double* fy, double* fz, double* f4, double* vsig, double* divv,
                     hipStream_t st);

} // namespace csph
