// cmgpu.hip -- MI355X (gfx950) implementation of the MDDF accumulation of
// ComplexMixtures.jl behind the C ABI of include/cmgpu.h.
//
// Two kernel families, chosen per job:
//
// "tiles" -- one solute molecule of any size (proteins, polymers), solvent
// molecules of up to 64 atoms.  Per batch of frames, on the job's stream:
//   k_radii             solvent molecule radius about its centre
//   k_grid              per-frame FP32 thresholds (guard band)
//   k_wrap_atoms, k_scan_cells, k_scatter
//                       solute atoms in a fine cell grid (counting sort + scan)
//   k_near_mark         cells within cutoff + radius of some solute atom
//   k_prep_solvent_tiles  centre of every solvent molecule, key when near
//   k_scan_rows, k_mol_scatter, k_chunk_table
//                       near molecules sorted by the Morton key of their cell
//                       per (frame, pass), cut into chunks of whole molecules,
//                       one wavefront each
//   k_fill_real         FP32 atoms of the near molecules in sorted order
//   k_eval_tiles        per chunk: AABB, solute atoms within the cutoff of it
//                       staged in LDS, FP32 minimum per atom (LDS broadcast),
//                       segmented minimum per molecule, exact FP64 resolve,
//                       LDS histograms
//   k_bulk_compact      ordered list of bulk molecules (inbulk, src/mddf.jl:55-57)
//   k_place_keys, k_place_near
//                       Philox placement (src/rigid_body.jl:107-137): centres
//                       of all slots, then the full placement of the slots whose
//                       centre is near the solute, in sorted order
//   then sort, chunk and k_eval_tiles again for the random passes.
//
// "mols" -- many solute molecules (self MDDF, small solutes): molecule
// centres in a cell list, one thread per solvent molecule evaluating all atom
// pairs of close molecule pairs (k_prep_solute_mols, k_prep_solvent,
// k_eval_mols_real, k_place, k_eval_mols_rand).
//
// k_flush turns integer hits into FP64 accumulators times the frame weight.
//
// Exactness.  Candidate pairs are evaluated in FP32 from coordinates in the
// frame of each molecule's cell.  Per atom (tiles) or molecule pair (mols) the
// kernels keep the best and second-best squared distance; when the gap exceeds
// a guard band that bounds the FP32 error, the FP32 argmin is the FP64 argmin
// and only that pair is recomputed in FP64 (the recorded distance).
// Otherwise every pair near the minimum is recomputed in FP64.  Records,
// indices and bins therefore equal those of an FP64 evaluation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cm_math.h"
#include "cmgpu.h"

using ull = unsigned long long;

namespace {

constexpr int BLOCK = 256;
constexpr int MAX_NBINS = 8192;

// ---------------------------------------------------------------------------
// device-side structures
// ---------------------------------------------------------------------------
struct Geo {                 // one frame of a batch
    cm::Cell64 cell;
    float H32[9];
    float Hi32[9];
    double sp[3];            // plane spacings
    int nc[3];
    int ncell;
    uint32_t frame;          // 1-based trajectory frame (keys the random streams)
    float cut2_hi;           // FP32 squared distances above this are beyond the cutoff
    float amb;               // best/second gap below which the FP64 resolve is exhaustive
    float rpre2;             // molecule-centre pre-check radius^2 (mols family)
    unsigned ru_bits, rv_bits;   // molecule radii (float bits, atomicMax)
    int err;                 // geometry the kernels cannot handle
    float rstage;            // staging radius around a chunk (cutoff + margin), tiles family
    float rnear;             // near-cell radius around a solute atom (cutoff + radius + margin)
    int kbits;               // bits per dimension of the Morton keys of the fine cells
    int ortho;               // diagonal cell matrix: cells are boxes (exact near-cell test)
    double guard;            // bound on |d2(FP32) - d2(FP64)| for pairs up to the cutoff
    float guard32;           // guard + (cutoff^2 + guard + 1) 2^-21, rounded up: covers FP32 b -+ guard32 rounding
};

struct KP {                  // job parameters, by value
    int Mu, nu, Mv, nv;
    int self_, iref, nbins, usecutoff;
    double binstep, cutsys, cut2, dbulk;
    float inv_binstep;       // 1 / binstep in FP32 (bin estimates only)
    // FP32 thresholds of the all-FP32 decision (decide32), rounded outward:
    // cutoff^2 down / up, binstep^2 (1 -+ 4e-7) down / up, dbulk^2 (1 -+ 4e-7)
    float c2lo32, c2hi32, bs2m32, bs2p32, db2lo32, db2hi32;
    int K;                   // random passes per frame (n_random_samples, 0 for coordination numbers)
    int kb;                  // bulk lists per frame (1 in the atoms family, K in the mols family)
    uint32_t key0, key1;
    const int *gu_ptr, *gu_idx, *gv_ptr, *gv_idx;
    int cap;                 // cell slots per frame
    int F;                   // frames in the batch
    int64_t Nu, Nv;
    int mpt;                 // molecules per chunk (tiles family)
    int mol_lane;            // tiles family, 3-atom solvent molecules: one molecule per lane row
    int tcap;                // staged solute atoms per round (tiles family)
    int glds;                // molecule-pair kernels: group counts privatised in LDS
    int gnu, gnv;            // group counts of the solute / solvent side
    int NP;                  // passes per frame, 1 + K (tiles family)
    int kcap;                // molecule sort keys per (frame, pass): Morton codes of the fine cells
    int place_all;           // molecule-pair family: place every slot of a random pass (verification)
};

// MinimumDistance record as the verification getters return it
// (src/minimum_distances.jl:12-19; 0-based i molecule-local, j global)
struct VRec {
    int within;
    int i;
    int64_t j;
    double d;
    int ref;
    int pad;
    double dref;
};

struct Bufs {
    Geo* geo;
    // cell list
    int* cnt;                // [F][cap+1]
    int* start;              // [F][cap+1]
    float4* tmp4;            // [F][nitems] unsorted entries
    int* tmpcell;            // [F][nitems]
    float4* sorted;          // [F][nitems]
    int nitems;              // atoms (atoms family) or solute molecules (mols family)
    // solute molecules (mols family)
    float4* solw;            // [F][Nu] atoms in the frame of their molecule's cell
    float4* ucen;            // [F][Mu]
    // solvent, real pass
    float4* vw;              // [F][Nv]
    int* vcell;              // [F][Mv]
    float4* vcen;            // [F][Mv]
    // random passes
    double* P;               // [F][K][Nv][3]
    float4* vw_r;            // [F][K][Nv]
    float4* vcen_r;          // [F][K][Mv]
    unsigned char* bflag;    // [F][kb][Mv]
    int* blist;              // [F][kb][Mv]
    int* nbulk;              // [F][kb]
    int* uk;                 // [F][K] drawn solute molecule of each random pass
    int* drawn_ptr;          // [F][Mu+1] CSR: solute molecule -> random passes
    int* drawn_idx;          // [F][K]
    // tiles family: near cells and molecules sorted by cell per (frame, pass)
    unsigned char* near;     // [F][cap]
    unsigned* dfield;        // [2][F][cap] 0xffffffff - bits of the distances from a near cell's centre and
                             // lower corner to the nearest solute atom image nearby (0: none), b_dfield
    int* mkey;               // [F][NP][Mv] Morton key of the centre's cell, -1 when far
    int* mrank;              // [F][NP][Mv] rank of a near molecule among those of its key (count order)
    int* mcnt;               // [F][NP][kcap+1]
    int* mstart;             // [F][NP][kcap+1]
    int* msorted;            // [F][NP][Mv]
    int* nnear;              // [F][NP]
    int* chunk_base;         // [F*npass + 2] chunk offsets of one launch, total, evaluation chunk counter
    // accumulators
    ull* cnt_md; ull* cnt_rdf; ull* cnt_md_r; ull* cnt_rdf_r;
    ull* cnt_gu; ull* cnt_gv; ull* cnt_gu_r; ull* cnt_gv_r;
    ull* stats;              // [0] exhaustive resolves, [2] geometry error bits of the job's frames
    VRec* vrec;              // verification: [F][NP][Mv] records of every pass (null when off)
};

// Device memory addressed as the global address space.  A pointer read from
// memory (the job table of the multi-job stages) is otherwise a generic one,
// and its loads and stores compile to flat instructions, which wait on the
// scalar-load counter too and take 64-bit VGPR addresses; through these they
// are global ones, and scalar loads where the address is uniform.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
typedef float f4v __attribute__((ext_vector_type(4)));

template <class T>
__device__ __forceinline__ T ldg(const T* p)
{
    return *(gptr<const T>)p;
}

// A frame's geometry by value, read through the constant address space (it
// is not written while the reading kernel runs): uniform addresses give
// scalar loads.  Not a template: the memcpy builtin is not available to the
// device side of function templates.
__device__ __forceinline__ Geo ldc_geo(const Geo* p)
{
    Geo g;
    __builtin_memcpy(&g, (__attribute__((address_space(4))) const Geo*)p, sizeof(Geo));
    return g;
}

struct Rec {
    int within;
    int i;
    int64_t j;
    double d;
    int ref;
    double dref;
    int bin;                 // 0-based bin of d when known without d (-1: from d)
    int rbin;                // the same for dref
    int bulk;                // d > dbulk when known without d (-1: from d)
};

__device__ __forceinline__ Rec rec_zero()
{
    Rec r;
    r.within = 0; r.i = -1; r.j = -1; r.d = INFINITY; r.ref = 0; r.dref = INFINITY;
    r.bin = -1; r.rbin = -1; r.bulk = -1;
    return r;
}

// offer a within-cutoff pair to an exact record: ties on d go to the
// smallest (i, j) (update_md, src/minimum_distances.jl:30-39)
__device__ __forceinline__ void rec_offer(Rec& r, int i, int64_t j, double d, bool isref)
{
    if (d < r.d || (d == r.d && (i < r.i || (i == r.i && j < r.j)))) {
        r.within = 1; r.i = i; r.j = j; r.d = d;
    }
    if (isref) { r.ref = 1; r.dref = fmin(r.dref, d); }
}

template <class T>
__device__ __forceinline__ void load3(const T* p, int64_t atom, double& x, double& y, double& z)
{
    x = (double)p[3 * atom]; y = (double)p[3 * atom + 1]; z = (double)p[3 * atom + 2];
}

__device__ __forceinline__ float4 mk4(float x, float y, float z, float w) { return make_float4(x, y, z, w); }

__device__ __forceinline__ void cell_of_frac(const Geo& g, double f0, double f1, double f2, int& cell, int& c0, int& c1,
                                             int& c2)
{
    c0 = (int)(f0 * g.nc[0]); if (c0 >= g.nc[0]) c0 = g.nc[0] - 1; if (c0 < 0) c0 = 0;
    c1 = (int)(f1 * g.nc[1]); if (c1 >= g.nc[1]) c1 = g.nc[1] - 1; if (c1 < 0) c1 = 0;
    c2 = (int)(f2 * g.nc[2]); if (c2 >= g.nc[2]) c2 = g.nc[2] - 1; if (c2 < 0) c2 = 0;
    cell = (c2 * g.nc[1] + c1) * g.nc[0] + c0;
}

__device__ __forceinline__ void cell_of_frac(const Geo& g, double f0, double f1, double f2, int& cell)
{
    int c0, c1, c2;
    cell_of_frac(g, f0, f1, f2, cell, c0, c1, c2);
}

// cell_of_frac of the fractional coordinates u01(r0), u01(r1), u01(r2), in
// integers: r 2^-32 is exact in FP64 and so is its product with nc (at most
// 32 + 21 significant bits), so (int)(u01(r) nc) = floor(r nc / 2^32) =
// (r nc) >> 32, always below nc
__device__ __forceinline__ void cell_of_draws(const Geo& g, uint32_t r0, uint32_t r1, uint32_t r2, int& cell, int& c0,
                                              int& c1, int& c2)
{
    c0 = (int)(((uint64_t)r0 * (uint64_t)g.nc[0]) >> 32);
    c1 = (int)(((uint64_t)r1 * (uint64_t)g.nc[1]) >> 32);
    c2 = (int)(((uint64_t)r2 * (uint64_t)g.nc[2]) >> 32);
    cell = (c2 * g.nc[1] + c1) * g.nc[0] + c0;
}

// fractional coordinates of a point, the lattice translation n = floor(f)
// that brings it into the cell, and the wrapped fractional part
__device__ __forceinline__ void wrap_point(const Geo& g, double x, double y, double z, double fr[3], double n[3])
{
    const double* Hi = g.cell.Hi;
    double f0 = Hi[0] * x + Hi[3] * y + Hi[6] * z;
    double f1 = Hi[1] * x + Hi[4] * y + Hi[7] * z;
    double f2 = Hi[2] * x + Hi[5] * y + Hi[8] * z;
    n[0] = floor(f0); n[1] = floor(f1); n[2] = floor(f2);
    fr[0] = f0 - n[0]; fr[1] = f1 - n[1]; fr[2] = f2 - n[2];
    for (int k = 0; k < 3; k++) if (fr[k] >= 1.0) { fr[k] = 0.0; n[k] += 1.0; }
}

// Molecule made whole around its first atom (minimum image of every atom
// relative to atom 0); returns the centre of coordinates.
template <class T>
__device__ void mol_centre(const Geo& g, const T* x, int64_t a0, int n, double c[3])
{
    double x0, y0, z0;
    load3(x, a0, x0, y0, z0);
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int a = 0; a < n; a++) {
        double xa, ya, za, ox, oy, oz;
        load3(x, a0 + a, xa, ya, za);
        cm::minimage(g.cell, xa - x0, ya - y0, za - z0, &ox, &oy, &oz);
        sx += ox; sy += oy; sz += oz;
    }
    c[0] = x0 + sx / n; c[1] = y0 + sy / n; c[2] = z0 + sz / n;
}

// ---------------------------------------------------------------------------
// radii and grid
// ---------------------------------------------------------------------------
// Radius bound of every molecule about its centre (molecule made whole around
// its first atom), FP32 with a margin that covers the FP32 rounding.
__device__ __forceinline__ void minimage32(const Geo& g, float dx, float dy, float dz, float& ox, float& oy, float& oz)
{
    const float* Hi = g.Hi32;
    const float* H = g.H32;
    float nx = rintf(Hi[0] * dx + Hi[3] * dy + Hi[6] * dz);
    float ny = rintf(Hi[1] * dx + Hi[4] * dy + Hi[7] * dz);
    float nz = rintf(Hi[2] * dx + Hi[5] * dy + Hi[8] * dz);
    ox = dx - (H[0] * nx + H[3] * ny + H[6] * nz);
    oy = dy - (H[1] * nx + H[4] * ny + H[7] * nz);
    oz = dz - (H[2] * nx + H[5] * ny + H[8] * nz);
}

// maximum over the threads of a block (blockDim.x a multiple of 64, at most 1024)
__device__ __forceinline__ float block_max(float x)
{
    __shared__ float part[16];
    for (int o = 32; o >= 1; o >>= 1) x = fmaxf(x, __shfl_xor(x, o, 64));
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[w] = x;
    __syncthreads();
    float r = part[0];
    for (int k = 1; k < nw; k++) r = fmaxf(r, part[k]);
    return r;
}

// molecule radii (bound on the distance of an atom from the centre of its
// molecule, with a margin over FP32 rounding); blockIdx.y = frame
template <class T>
__device__ __forceinline__ void b_radii(KP P, Bufs B, const T* xu, const T* xv, int do_solute, int bx_, int nbx_, int by_, int nby_)
{
    const int f = by_;
    Geo& g = B.geo[f];
    const int64_t nmol = (int64_t)P.Mv + (do_solute ? P.Mu : 0);
    float rv = 0.f, ru = 0.f;
    for (int64_t m = bx_ * (int64_t)blockDim.x + threadIdx.x; m < nmol; m += (int64_t)nbx_ * blockDim.x) {
        const bool solv = m < P.Mv;
        const T* x = solv ? xv + (int64_t)f * P.Nv * 3 : xu + (int64_t)f * P.Nu * 3;
        const int n = solv ? P.nv : P.nu;
        const int64_t a0 = solv ? m * P.nv : (m - P.Mv) * P.nu;
        double x0, y0, z0;
        load3(x, a0, x0, y0, z0);
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (int a = 0; a < n; a++) {
            double xa, ya, za;
            load3(x, a0 + a, xa, ya, za);
            float ox, oy, oz;
            minimage32(g, (float)(xa - x0), (float)(ya - y0), (float)(za - z0), ox, oy, oz);
            sx += ox; sy += oy; sz += oz;
        }
        sx /= n; sy /= n; sz /= n;
        float r2 = 0.f;
        for (int a = 0; a < n; a++) {
            double xa, ya, za;
            load3(x, a0 + a, xa, ya, za);
            float ox, oy, oz;
            minimage32(g, (float)(xa - x0), (float)(ya - y0), (float)(za - z0), ox, oy, oz);
            float dx = ox - sx, dy = oy - sy, dz = oz - sz;
            r2 = fmaxf(r2, dx * dx + dy * dy + dz * dz);
        }
        const float r = sqrtf(r2) * 1.0001f + 1e-3f;
        if (solv) rv = fmaxf(rv, r); else ru = fmaxf(ru, r);
    }
    rv = block_max(rv);
    ru = block_max(ru);
    if (threadIdx.x == 0) {
        if (rv > 0.f) atomicMax(&g.rv_bits, __float_as_uint(rv));
        if (ru > 0.f) atomicMax(&g.ru_bits, __float_as_uint(ru));
    }
}

template <class T>
__global__ void k_radii(KP P, Bufs B, const T* xu, const T* xv, int do_solute)
{
    b_radii<T>(P, B, xu, xv, do_solute, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

__device__ __forceinline__ double halfulp32(double x)
{
    if (!(x > 1e-30)) return 0x1p-150;
    return ldexp(1.0, ilogb(x) - 24);
}

// geometry errors of a batch's frames into the job's error bits (bit 0:
// molecules too large for the cell, bit 1: grid larger than its allocation),
// read with the accumulators at the end of the job
__device__ __forceinline__ void b_geo_flags(KP P, Bufs B, int bx_, int nbx_, int by_, int nby_)
{
    for (int f = threadIdx.x; f < P.F; f += blockDim.x) {
        const int e = B.geo[f].err;
        if (e) atomicOr(&B.stats[2], e == 2 ? 1ull : 2ull);
    }
}

__global__ void k_geo_flags(KP P, Bufs B)
{
    b_geo_flags(P, B, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

__device__ __forceinline__ void b_grid(KP P, Bufs B, int mols_family, int bx_, int nbx_, int by_, int nby_)
{
    int f = bx_ * blockDim.x + threadIdx.x;
    if (f >= P.F) return;
    Geo& g = B.geo[f];
    double ru = mols_family ? (double)__uint_as_float(g.ru_bits) : 0.0;
    double rv = (double)__uint_as_float(g.rv_bits);
    if (mols_family) {
        double side = P.cutsys + ru + rv + 1e-3 + 1e-7 * (g.sp[0] + g.sp[1] + g.sp[2]);
        int ncell = 1;
        for (int k = 0; k < 3; k++) {
            int n = (int)floor(g.sp[k] / side);
            g.nc[k] = n < 1 ? 1 : n;
            ncell *= g.nc[k];
            if (!(side < g.sp[k])) g.err = 2;   // image enumeration assumes ru+rv+cutoff < spacing
        }
        if (ncell > P.cap) g.err = 3;
        g.ncell = ncell;
    }
    const double* H = g.cell.H;
    double xm = 0.0;
    for (int r = 0; r < 3; r++) xm = fmax(xm, fabs(H[r]) + fabs(H[3 + r]) + fabs(H[6 + r]));
    xm += fmax(ru, rv) + P.cutsys + 1.0;
    double E = 10.0 * halfulp32(2.0 * xm) + 1e-7;
    double dmax = P.cutsys + 0.5;
    double G = E * (2.0 * 1.7320508075688772 * dmax + 3.0 * E) + 8.0 * 0x1p-24 * dmax * dmax;
    g.guard = G;
    g.guard32 = nextafterf((float)(G + (P.cut2 + G + 1.0) * 0x1p-21), INFINITY);
    float hi = (float)(P.cut2 + G);
    g.cut2_hi = nextafterf(hi, INFINITY);
    float ulpc = (float)(halfulp32((double)g.cut2_hi) * 2.0);
    g.amb = nextafterf((float)(2.0 * G) + 8.0f * ulpc, INFINITY);
    float rp = (float)(P.cutsys + ru + rv + 1e-2);
    g.rpre2 = rp * rp;
    g.rstage = (float)(P.cutsys + 20.0 * E + 1e-3);
    g.rnear = (float)(P.cutsys + rv + 20.0 * E + 1e-2);
}

__global__ void k_grid(KP P, Bufs B, int mols_family)
{
    b_grid(P, B, mols_family, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// ---------------------------------------------------------------------------
// cell list: counting sort by cell
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void b_wrap_atoms(KP P, Bufs B, const T* xu, int bx_, int nbx_, int by_, int nby_)
{
    int64_t total = (int64_t)P.F * P.nu;
    for (int64_t it = bx_ * (int64_t)blockDim.x + threadIdx.x; it < total; it += (int64_t)nbx_ * blockDim.x) {
        int f = (int)(it / P.nu);
        int i = (int)(it - (int64_t)f * P.nu);
        const Geo g = ldc_geo(B.geo + f);
        double x, y, z;
        load3(xu + (int64_t)f * P.Nu * 3, i, x, y, z);
        double fr[3], n[3];
        wrap_point(g, x, y, z, fr, n);
        int cell;
        cell_of_frac(g, fr[0], fr[1], fr[2], cell);
        const double* H = g.cell.H;
        double wx = H[0] * fr[0] + H[3] * fr[1] + H[6] * fr[2];
        double wy = H[1] * fr[0] + H[4] * fr[1] + H[7] * fr[2];
        double wz = H[2] * fr[0] + H[5] * fr[1] + H[8] * fr[2];
        int64_t o = (int64_t)f * B.nitems + i;
        B.tmp4[o] = mk4((float)wx, (float)wy, (float)wz, __int_as_float(i));
        B.tmpcell[o] = cell;
        atomicAdd(&B.cnt[(int64_t)f * (P.cap + 1) + cell], 1);
    }
}

template <class T>
__global__ void k_wrap_atoms(KP P, Bufs B, const T* xu)
{
    b_wrap_atoms<T>(P, B, xu, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// Solute molecules (mols family): made whole, centre wrapped into the cell,
// atoms written in the frame of that cell.
template <class T>
__global__ void k_prep_solute_mols(KP P, Bufs B, const T* xu)
{
    int64_t total = (int64_t)P.F * P.Mu;
    for (int64_t it = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; it < total; it += (int64_t)gridDim.x * blockDim.x) {
        int f = (int)(it / P.Mu);
        int u = (int)(it - (int64_t)f * P.Mu);
        const Geo& g = B.geo[f];
        const T* x = xu + (int64_t)f * P.Nu * 3;
        int64_t a0 = (int64_t)u * P.nu;
        double c[3];
        mol_centre(g, x, a0, P.nu, c);
        double fr[3], n[3];
        wrap_point(g, c[0], c[1], c[2], fr, n);
        int cell;
        cell_of_frac(g, fr[0], fr[1], fr[2], cell);
        const double* H = g.cell.H;
        double tx = H[0] * n[0] + H[3] * n[1] + H[6] * n[2];
        double ty = H[1] * n[0] + H[4] * n[1] + H[7] * n[2];
        double tz = H[2] * n[0] + H[5] * n[1] + H[8] * n[2];
        double x0, y0, z0;
        load3(x, a0, x0, y0, z0);
        for (int a = 0; a < P.nu; a++) {
            double xa, ya, za, ox, oy, oz;
            load3(x, a0 + a, xa, ya, za);
            cm::minimage(g.cell, xa - x0, ya - y0, za - z0, &ox, &oy, &oz);
            B.solw[(int64_t)f * P.Nu + a0 + a] = mk4((float)(x0 + ox - tx), (float)(y0 + oy - ty), (float)(z0 + oz - tz), 0.f);
        }
        float4 cw = mk4((float)(c[0] - tx), (float)(c[1] - ty), (float)(c[2] - tz), __int_as_float(u));
        B.ucen[(int64_t)f * P.Mu + u] = cw;
        int64_t o = (int64_t)f * B.nitems + u;
        B.tmp4[o] = cw;
        B.tmpcell[o] = cell;
        atomicAdd(&B.cnt[(int64_t)f * (P.cap + 1) + cell], 1);
    }
}

// exclusive scan of the cell counts, one block per frame; leaves cnt zeroed
__device__ __forceinline__ void b_scan_cells(KP P, Bufs B, int bx_, int nbx_, int by_, int nby_)
{
    __shared__ int wsum[16];
    __shared__ int carry;
    int f = bx_;
    int ncell = B.geo[f].ncell;
    int* cnt = B.cnt + (int64_t)f * (P.cap + 1);
    int* st = B.start + (int64_t)f * (P.cap + 1);
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int base = 0; base < ncell; base += blockDim.x) {
        int idx = base + threadIdx.x;
        int v = idx < ncell ? cnt[idx] : 0;
        int incl = v;
        for (int o = 1; o < 64; o <<= 1) {
            int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int woff = 0;
        for (int w = 0; w < wid; w++) woff += wsum[w];
        int tot = 0;
        for (int w = 0; w < nw; w++) tot += wsum[w];
        if (idx < ncell) { st[idx] = carry + woff + incl - v; cnt[idx] = 0; }
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) st[ncell] = carry;
}

__global__ void __launch_bounds__(1024) k_scan_cells(KP P, Bufs B)
{
    b_scan_cells(P, B, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

__device__ __forceinline__ void b_scatter(KP P, Bufs B, int nper, int bx_, int nbx_, int by_, int nby_)
{
    int64_t total = (int64_t)P.F * nper;
    for (int64_t it = bx_ * (int64_t)blockDim.x + threadIdx.x; it < total; it += (int64_t)nbx_ * blockDim.x) {
        int f = (int)(it / nper);
        int e = (int)(it - (int64_t)f * nper);
        int64_t o = (int64_t)f * B.nitems + e;
        int cell = B.tmpcell[o];
        int64_t cb = (int64_t)f * (P.cap + 1);
        int pos = B.start[cb + cell] + atomicAdd(&B.cnt[cb + cell], 1);
        B.sorted[(int64_t)f * B.nitems + pos] = B.tmp4[o];
    }
}

__global__ void k_scatter(KP P, Bufs B, int nper)
{
    b_scatter(P, B, nper, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// solvent molecules of the real pass, in the frame of their cell
template <class T>
__global__ void k_prep_solvent(KP P, Bufs B, const T* xv)
{
    int64_t total = (int64_t)P.F * P.Mv;
    for (int64_t it = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; it < total; it += (int64_t)gridDim.x * blockDim.x) {
        int f = (int)(it / P.Mv);
        int v = (int)(it - (int64_t)f * P.Mv);
        const Geo& g = B.geo[f];
        const T* x = xv + (int64_t)f * P.Nv * 3;
        int64_t a0 = (int64_t)v * P.nv;
        double c[3];
        mol_centre(g, x, a0, P.nv, c);
        double fr[3], n[3];
        wrap_point(g, c[0], c[1], c[2], fr, n);
        int cell;
        cell_of_frac(g, fr[0], fr[1], fr[2], cell);
        const double* H = g.cell.H;
        double tx = H[0] * n[0] + H[3] * n[1] + H[6] * n[2];
        double ty = H[1] * n[0] + H[4] * n[1] + H[7] * n[2];
        double tz = H[2] * n[0] + H[5] * n[1] + H[8] * n[2];
        double x0, y0, z0;
        load3(x, a0, x0, y0, z0);
        B.vcell[(int64_t)f * P.Mv + v] = cell;
        if (P.self_) continue;   // the pair kernel reads the solute copies (k_prep_solute_mols)
        for (int a = 0; a < P.nv; a++) {
            double xa, ya, za, ox, oy, oz;
            load3(x, a0 + a, xa, ya, za);
            cm::minimage(g.cell, xa - x0, ya - y0, za - z0, &ox, &oy, &oz);
            B.vw[(int64_t)f * P.Nv + a0 + a] = mk4((float)(x0 + ox - tx), (float)(y0 + oy - ty), (float)(z0 + oz - tz), 0.f);
        }
        B.vcen[(int64_t)f * P.Mv + v] = mk4((float)(c[0] - tx), (float)(c[1] - ty), (float)(c[2] - tz), 0.f);
    }
}

// ---------------------------------------------------------------------------
// histogram updates (update_counters!, src/update_counters.jl:43-88)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void group_hit(ull* gc, int nbins, int b, int64_t atom, const int* ptr, const int* idx, int natomspermol)
{
    if (!ptr) {
        int64_t t = atom % natomspermol;
        atomicAdd(&gc[t * nbins + b], 1ull);
    } else {
        for (int k = ptr[atom]; k < ptr[atom + 1]; k++) atomicAdd(&gc[(int64_t)idx[k] * nbins + b], 1ull);
    }
}

// bins and the dbulk test of a record from its distances where the fast
// decision did not set them (setbin, src/results.jl:28)
__device__ __forceinline__ void rec_complete(const KP& P, Rec& r)
{
    if (r.within && r.bin < 0) {
        r.bin = (int)cm::setbin(r.d, P.binstep) - 1;
        r.bulk = r.d > P.dbulk ? 1 : 0;
    }
    if (r.ref && r.rbin < 0) r.rbin = (int)cm::setbin(r.dref, P.binstep) - 1;
}

__device__ __forceinline__ void hist_record(const KP& P, unsigned* sh, ull* gu, ull* gv, const Rec& r)
{
    if (!r.within) return;
    int b = r.bin;
    if (b >= P.nbins) return;   // unreachable: d <= cutoff (validated at job creation)
    atomicAdd(&sh[b], 1u);
    group_hit(gu, P.nbins, b, r.i, P.gu_ptr, P.gu_idx, P.nu);
    if (P.self_) group_hit(gu, P.nbins, b, r.j, P.gv_ptr, P.gv_idx, P.nu);
    else group_hit(gv, P.nbins, b, r.j, P.gv_ptr, P.gv_idx, P.nv);
    if (r.ref) {
        const int rb = r.rbin;
        if (rb < P.nbins) atomicAdd(&sh[P.nbins + rb], 1u);
    }
}

// the same with the md/rdf counts straight into the global counters
__device__ __forceinline__ void hist_record_global(const KP& P, ull* md, ull* rdf, ull* gu, ull* gv, const Rec& r)
{
    if (!r.within) return;
    const int b = r.bin;
    if (b >= P.nbins) return;   // unreachable: d <= cutoff (validated at job creation)
    atomicAdd(&md[b], 1ull);
    group_hit(gu, P.nbins, b, r.i, P.gu_ptr, P.gu_idx, P.nu);
    if (P.self_) group_hit(gu, P.nbins, b, r.j, P.gv_ptr, P.gv_idx, P.nu);
    else group_hit(gv, P.nbins, b, r.j, P.gv_ptr, P.gv_idx, P.nv);
    if (r.ref) {
        const int rb = r.rbin;
        if (rb < P.nbins) atomicAdd(&rdf[rb], 1ull);
    }
}

// Group counts privatised in LDS (small group arrays): the same updates as
// group_hit into a block-local copy, flushed once per block.
__device__ __forceinline__ void group_hit_lds(unsigned* gc, int nbins, int b, int64_t atom, const int* ptr, const int* idx,
                                              int natomspermol)
{
    if (!ptr) {
        const int64_t t = atom % natomspermol;
        atomicAdd(&gc[t * nbins + b], 1u);
    } else {
        for (int k = ptr[atom]; k < ptr[atom + 1]; k++) atomicAdd(&gc[(int64_t)idx[k] * nbins + b], 1u);
    }
}

// md/rdf into the block histogram sh; groups into LDS (shg) when privatised,
// else straight into the global counters
__device__ __forceinline__ void hist_record_mols(const KP& P, unsigned* sh, unsigned* shg, ull* gu, ull* gv, const Rec& r)
{
    if (!shg) {
        hist_record(P, sh, gu, gv, r);
        return;
    }
    if (!r.within) return;
    const int b = r.bin;
    if (b >= P.nbins) return;
    atomicAdd(&sh[b], 1u);
    unsigned* shu = shg;
    unsigned* shv = shg + (int64_t)P.nbins * P.gnu;
    group_hit_lds(shu, P.nbins, b, r.i, P.gu_ptr, P.gu_idx, P.nu);
    if (P.self_) group_hit_lds(shu, P.nbins, b, r.j, P.gv_ptr, P.gv_idx, P.nu);
    else group_hit_lds(shv, P.nbins, b, r.j, P.gv_ptr, P.gv_idx, P.nv);
    if (r.ref) {
        const int rb = r.rbin;
        if (rb < P.nbins) atomicAdd(&sh[P.nbins + rb], 1u);
    }
}

__device__ __forceinline__ void groups_flush(const KP& P, const unsigned* shg, ull* gu, ull* gv)
{
    __syncthreads();
    const int nu_ = P.gnu * P.nbins, nv_ = P.self_ ? 0 : P.gnv * P.nbins;
    for (int k = threadIdx.x; k < nu_; k += blockDim.x)
        if (shg[k]) atomicAdd(&gu[k], (ull)shg[k]);
    for (int k = threadIdx.x; k < nv_; k += blockDim.x)
        if (shg[nu_ + k]) atomicAdd(&gv[k], (ull)shg[nu_ + k]);
}

__device__ __forceinline__ void hist_clear(unsigned* sh, int n)
{
    for (int k = threadIdx.x; k < n; k += blockDim.x) sh[k] = 0u;
    __syncthreads();
}

__device__ __forceinline__ void hist_flush(unsigned* sh, int nbins, ull* md, ull* rdf)
{
    __syncthreads();
    for (int k = threadIdx.x; k < nbins; k += blockDim.x) {
        if (sh[k]) atomicAdd(&md[k], (ull)sh[k]);
        if (sh[nbins + k]) atomicAdd(&rdf[k], (ull)sh[nbins + k]);
    }
}

__device__ __forceinline__ bool inbulk(const KP& P, const Rec& r)
{
    return P.usecutoff ? (r.within && r.bulk == 1) : !r.within;
}

// A record from the FP32 minimum b of a unique argmin, without the FP64
// recompute, when every FP64 d2 in [b - G, b + G] gives the same cutoff test,
// bin and dbulk test.  The bin of d2 is max(1, ceil(fl(sqrt(d2)) / binstep));
// between ((k-1) binstep)^2 (1 + 1e-12) and (k binstep)^2 (1 - 1e-12) it is k
// whatever the FP64 rounding of the square root and the quotient (below 1e-15
// relative), and likewise d > dbulk is decided outside dbulk^2 (1 -+ 1e-12).
// All in FP32 and branch-free: [lo, hi] = b -+ guard32 (guard32 covers the
// FP32 rounding of both sums, so [lo, hi] contains [b - G, b + G]), and every
// threshold is rounded outward: the cutoff^2 bounds, k^2 binstep^2 (1 - 4e-7)
// <= (k binstep)^2 (1 - 1e-12) and (k-1)^2 binstep^2 (1 + 4e-7) >= ((k-1)
// binstep)^2 (1 + 1e-12) after the FP32 roundings of k^2 and the products
// (each below 1.2e-7 relative), and the dbulk^2 bounds likewise.  The bin
// estimate k comes from the hardware square root; any k is safe, the edge
// tests accept it only when it is the bin of every d2 in [lo, hi].  Returns 1
// when no pair can be within the cutoff, 2 when the record is within with the
// 0-based bin and the dbulk flag set, 0 when the interval straddles a decision.
__device__ __forceinline__ int decide32(const KP& P, float G32, float b, int& bin, int& bulk)
{
    const float lo = b - G32, hi = b + G32;
    const int k = max((int)ceilf(__builtin_amdgcn_sqrtf(b) * P.inv_binstep), 1);
    const float kf = (float)k, km = kf - 1.0f;
    const bool inbin = hi < kf * kf * P.bs2m32 && (k == 1 || lo > km * km * P.bs2p32);
    const bool b1 = lo > P.db2hi32, b0 = hi < P.db2lo32;
    const bool okbulk = !P.usecutoff || b1 || b0;
    bin = k - 1;
    bulk = (P.usecutoff && b1) ? 1 : 0;
    const bool within = hi <= P.c2lo32 && inbin && okbulk;
    return lo > P.c2hi32 ? 1 : (within ? 2 : 0);
}

// FP32 running minimum with second best
struct Best {
    float best, second;
    int bi, ba;
    __device__ __forceinline__ void init() { best = INFINITY; second = INFINITY; bi = -1; ba = -1; }
    // branch-free: second smallest = median of (best, second, d2) while
    // best <= second; selects rather than a branch keep the loop straight
    __device__ __forceinline__ void offer(float d2, int i, int a)
    {
        const bool lt = d2 < best;
        second = __builtin_amdgcn_fmed3f(best, second, d2);
        best = lt ? d2 : best;
        bi = lt ? i : bi;
        ba = lt ? a : ba;
    }
};

// neighbour cell (with its lattice shift) of cell c at offset o in one dimension
__device__ __forceinline__ int nb_cell(int c, int o, int n, int& m)
{
    int t = c + o;
    m = t < 0 ? -1 : (t >= n ? 1 : 0);
    return t - m * n;
}

// Wavefront scans and reductions with DPP row shifts and row broadcasts
// (row_shr:1,2,4,8 within rows of 16 lanes, then row_bcast:15 into rows 1
// and 3 and row_bcast:31 into rows 2 and 3); lanes without a source read the
// identity.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int identity, int x)
{
    return __builtin_amdgcn_update_dpp(identity, x, CTRL, ROWS, 0xf, false);
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += dpp_i<0x111, 0xf>(0, v);
    v += dpp_i<0x112, 0xf>(0, v);
    v += dpp_i<0x114, 0xf>(0, v);
    v += dpp_i<0x118, 0xf>(0, v);
    v += dpp_i<0x142, 0xa>(0, v);
    v += dpp_i<0x143, 0xc>(0, v);
    return v;
}

__device__ __forceinline__ float fmin_sel(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmax_sel(float a, float b) { return b > a ? b : a; }

// a wave-uniform float moved to a scalar register
__device__ __forceinline__ float sgpr_f(float x)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// minimum / maximum over the wavefront, the same in every lane
__device__ __forceinline__ float wave_min(float x)
{
    const int id = __float_as_int(INFINITY);
    x = fmin_sel(x, __int_as_float(dpp_i<0x111, 0xf>(id, __float_as_int(x))));
    x = fmin_sel(x, __int_as_float(dpp_i<0x112, 0xf>(id, __float_as_int(x))));
    x = fmin_sel(x, __int_as_float(dpp_i<0x114, 0xf>(id, __float_as_int(x))));
    x = fmin_sel(x, __int_as_float(dpp_i<0x118, 0xf>(id, __float_as_int(x))));
    x = fmin_sel(x, __int_as_float(dpp_i<0x142, 0xa>(id, __float_as_int(x))));
    x = fmin_sel(x, __int_as_float(dpp_i<0x143, 0xc>(id, __float_as_int(x))));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}
__device__ __forceinline__ float wave_max(float x)
{
    const int id = __float_as_int(-INFINITY);
    x = fmax_sel(x, __int_as_float(dpp_i<0x111, 0xf>(id, __float_as_int(x))));
    x = fmax_sel(x, __int_as_float(dpp_i<0x112, 0xf>(id, __float_as_int(x))));
    x = fmax_sel(x, __int_as_float(dpp_i<0x114, 0xf>(id, __float_as_int(x))));
    x = fmax_sel(x, __int_as_float(dpp_i<0x118, 0xf>(id, __float_as_int(x))));
    x = fmax_sel(x, __int_as_float(dpp_i<0x142, 0xa>(id, __float_as_int(x))));
    x = fmax_sel(x, __int_as_float(dpp_i<0x143, 0xc>(id, __float_as_int(x))));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// The frame fields the pair kernels read on every pair, copied into registers
// once: read through the Geo reference they would be reloaded from memory after
// every global store or atomic (which may alias the geometry array).
struct GeoHot {
    float H32[9];
    float cut2_hi, amb, rpre2, guard32;
    double guard;
    int nc[3];
    __device__ explicit GeoHot(const Geo& g)
        : cut2_hi(g.cut2_hi), amb(g.amb), rpre2(g.rpre2), guard32(g.guard32), guard(g.guard)
    {
#pragma unroll
        for (int k = 0; k < 9; k++) H32[k] = g.H32[k];
        nc[0] = g.nc[0]; nc[1] = g.nc[1]; nc[2] = g.nc[2];
    }
};

template <class GEO>
__device__ __forceinline__ void shift_of(const GEO& g, int mx, int my, int mz, float& tx, float& ty, float& tz)
{
    float fx = (float)mx, fy = (float)my, fz = (float)mz;
    tx = g.H32[0] * fx + g.H32[3] * fy + g.H32[6] * fz;
    ty = g.H32[1] * fx + g.H32[4] * fy + g.H32[7] * fz;
    tz = g.H32[2] * fx + g.H32[5] * fy + g.H32[8] * fz;
}

// ---------------------------------------------------------------------------
// minimum distances, mols family (many solute molecules)
// ---------------------------------------------------------------------------
// FP32 evaluation of all atom pairs of solute molecule u (atoms xs) and the
// solvent molecule whose atoms are yw, the solute shifted by t.
__device__ __forceinline__ void pairs_mol(const KP& P, const float4* xs, const float4* yw, float tx, float ty, float tz,
                                          Best& bm, Best& br)
{
    for (int a = 0; a < P.nv; a++) {
        float4 y = yw[a];
        float yx = y.x - tx, yy = y.y - ty, yz = y.z - tz;
        bool isref = (a == P.iref);
        for (int i = 0; i < P.nu; i++) {
            float4 s = xs[i];
            float dx = yx - s.x, dy = yy - s.y, dz = yz - s.z;
            float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            bm.offer(d2, i, a);
            if (isref) br.offer(d2, i, a);
        }
    }
}

// exact record of the molecule pair (u, v) from the FP32 bests
template <class TU, class YS>
__device__ Rec resolve_mols(const KP& P, const Geo& g, const TU* xuf, int u, int v, YS ysrc, const Best& bm, const Best& br,
                            ull* stats)
{
    Rec r = rec_zero();
    bool exhaustive = false;
    int64_t ibase = (int64_t)u * P.nu, jbase = (int64_t)v * P.nv;
    if (bm.best <= g.cut2_hi) {
        if (bm.second - bm.best > g.amb) {
            int bn, bk;
            const int dd = decide32(P, g.guard32, bm.best, bn, bk);
            if (dd != 0) {
                if (dd == 2) { r.within = 1; r.i = bm.bi; r.j = jbase + bm.ba; r.d = bm.best; r.bin = bn; r.bulk = bk; }
            } else {
                double x0, x1, x2, y0, y1, y2;
                load3(xuf, ibase + bm.bi, x0, x1, x2);
                ysrc(bm.ba, y0, y1, y2);
                double d2 = cm::canon_d2(g.cell, x0, x1, x2, y0, y1, y2);
                if (d2 <= P.cut2) { r.within = 1; r.i = bm.bi; r.j = jbase + bm.ba; r.d = sqrt(d2); }
            }
        } else exhaustive = true;
    }
    if (br.best <= g.cut2_hi) {
        if (br.second - br.best > g.amb) {
            int bn, bk;
            const int dd = decide32(P, g.guard32, br.best, bn, bk);
            if (dd != 0) {
                if (dd == 2) { r.ref = 1; r.dref = br.best; r.rbin = bn; }
            } else {
                double x0, x1, x2, y0, y1, y2;
                load3(xuf, ibase + br.bi, x0, x1, x2);
                ysrc(P.iref, y0, y1, y2);
                double d2 = cm::canon_d2(g.cell, x0, x1, x2, y0, y1, y2);
                if (d2 <= P.cut2) { r.ref = 1; r.dref = sqrt(d2); }
            }
        } else exhaustive = true;
    }
    if (exhaustive) {
        atomicAdd(stats, 1ull);
        r = rec_zero();
        for (int a = 0; a < P.nv; a++) {
            double y0, y1, y2;
            ysrc(a, y0, y1, y2);
            for (int i = 0; i < P.nu; i++) {
                double x0, x1, x2;
                load3(xuf, ibase + i, x0, x1, x2);
                double d2 = cm::canon_d2(g.cell, x0, x1, x2, y0, y1, y2);
                if (d2 <= P.cut2) rec_offer(r, i, jbase + a, sqrt(d2), a == P.iref);
            }
        }
    }
    rec_complete(P, r);
    return r;
}

// LDS pair queue of the molecule-pair kernel: entries per wavefront
constexpr int QCAP = 128;
constexpr int DCAP = 128;   // deferred FP64 entries per wavefront (self)
constexpr int GLDS_MAX = 32 * 1024;   // bytes of LDS group counts per block

__device__ __forceinline__ int wave_max_i(int x)
{
    x = max(x, dpp_i<0x111, 0xf>(INT_MIN, x));
    x = max(x, dpp_i<0x112, 0xf>(INT_MIN, x));
    x = max(x, dpp_i<0x114, 0xf>(INT_MIN, x));
    x = max(x, dpp_i<0x118, 0xf>(INT_MIN, x));
    x = max(x, dpp_i<0x142, 0xa>(INT_MIN, x));
    x = max(x, dpp_i<0x143, 0xc>(INT_MIN, x));
    return __builtin_amdgcn_readlane(x, 63);
}

// Lattice shifts m (per dimension, at most three) that map the cell index cv
// of one molecule onto the neighbour cell cu of another, for the offsets o in
// {-1, 0, 1} with wrap(cv + o) == cu, as a set: bit m + 1.  With three or
// more cells per dimension there is exactly one; small grids map several
// offsets onto one cell.  (A bit set rather than an array: dynamically
// indexed arrays live in scratch memory.)
__device__ __forceinline__ int shifts_onto(int cv, int cu, int n)
{
    int set = 0;
    for (int o = -1; o <= 1; o++) {
        int mm;
        if (nb_cell(cv, o, n, mm) == cu) set |= 1 << (mm + 1);
    }
    return set;
}

// the k-th (from 0, ascending) lattice shift of a shift set
__device__ __forceinline__ int nth_shift(int set, int k)
{
    if (k > 0) set &= set - 1;
    if (k > 1) set &= set - 1;
    return __builtin_ctz((unsigned)set) - 1;
}

// Per-frame state of the molecule-pair kernels: records of molecule pairs
// from their FP32 bests (resolve, histograms, bulk flags) and the evaluation
// of one queued pair.
template <class TU, class TV, int NU>
struct PairCtx {
    const KP& P;
    const Bufs& B;
    const Geo& g;
    const GeoHot gh;
    int f;
    const float4* solw;
    const float4* vwf;
    const float4* vcen;
    const int* vcellf;
    const TU* xuf;
    const TV* xvf;
    const int* dp;
    const int* di;
    unsigned* sh;
    unsigned* shg;

    __device__ PairCtx(const KP& P_, const Bufs& B_, int f_, const TU* xu, const TV* xv, unsigned* sh_, unsigned* shg_)
        : P(P_), B(B_), g(B_.geo[f_]), gh(B_.geo[f_]), f(f_), solw(B_.solw + (int64_t)f_ * P_.Nu),
          // self: the solvent side reads the solute copies (the same molecules,
          // wrapped identically), which halves the data each frame touches
          vwf(P_.self_ ? B_.solw + (int64_t)f_ * P_.Nu : B_.vw + (int64_t)f_ * P_.Nv),
          vcen(P_.self_ ? B_.ucen + (int64_t)f_ * P_.Mu : B_.vcen + (int64_t)f_ * P_.Mv), vcellf(B_.vcell + (int64_t)f_ * P_.Mv), xuf(xu + (int64_t)f_ * P_.Nu * 3),
          xvf(xv + (int64_t)f_ * P_.Nv * 3), dp(B_.drawn_ptr + (int64_t)f_ * (P_.Mu + 1)), di(B_.drawn_idx + (int64_t)f_ * P_.K),
          sh(sh_), shg(shg_) {}

    // record of solute molecule uu and solvent molecule vv
    __device__ void emit(int uu, int vv, const Best& bm, const Best& br) const
    {
        const int64_t jb = (int64_t)vv * P.nv;
        const TV* xv = xvf;
        auto ysrc = [&](int a, double& y0, double& y1, double& y2) { load3(xv, jb + a, y0, y1, y2); };
        Rec r = resolve_mols(P, g, xuf, uu, vv, ysrc, bm, br, &B.stats[0]);
        hist_record_mols(P, sh, shg, B.cnt_gu, B.cnt_gv, r);
        if (P.K > 0 && r.within)
            for (int qq = dp[uu]; qq < dp[uu + 1]; qq++)
                B.bflag[((int64_t)f * P.kb + di[qq]) * P.Mv + vv] = inbulk(P, r) ? 1 : 0;
    }

    // md (+2), group and dbulk-flag updates of both orientations of the self
    // pair {u, v} whose minimum distance joins atom bi of u and atom ba of v,
    // in 0-based bin mbin (mbulk: d > dbulk)
    __device__ void self_main(int u, int v, int bi, int ba, int mbin, int mbulk) const
    {
        if (mbin >= P.nbins) return;   // unreachable: d <= cutoff
        atomicAdd(&sh[mbin], 2u);
        if (!P.gu_ptr && !P.gv_ptr) {
            if (shg) {
                atomicAdd(&shg[(int64_t)bi * P.nbins + mbin], 2u);
                atomicAdd(&shg[(int64_t)ba * P.nbins + mbin], 2u);
            } else {
                atomicAdd(&B.cnt_gu[(int64_t)bi * P.nbins + mbin], 2ull);
                atomicAdd(&B.cnt_gu[(int64_t)ba * P.nbins + mbin], 2ull);
            }
        } else {
            // custom groups: the solute side is looked up with the
            // molecule-local index, the solvent side with the global one
            const int64_t ju = (int64_t)u * P.nv + bi, jv = (int64_t)v * P.nv + ba;
            if (shg) {
                group_hit_lds(shg, P.nbins, mbin, bi, P.gu_ptr, P.gu_idx, P.nu);
                group_hit_lds(shg, P.nbins, mbin, jv, P.gv_ptr, P.gv_idx, P.nu);
                group_hit_lds(shg, P.nbins, mbin, ba, P.gu_ptr, P.gu_idx, P.nu);
                group_hit_lds(shg, P.nbins, mbin, ju, P.gv_ptr, P.gv_idx, P.nu);
            } else {
                group_hit(B.cnt_gu, P.nbins, mbin, bi, P.gu_ptr, P.gu_idx, P.nu);
                group_hit(B.cnt_gu, P.nbins, mbin, jv, P.gv_ptr, P.gv_idx, P.nu);
                group_hit(B.cnt_gu, P.nbins, mbin, ba, P.gu_ptr, P.gu_idx, P.nu);
                group_hit(B.cnt_gu, P.nbins, mbin, ju, P.gv_ptr, P.gv_idx, P.nu);
            }
        }
        if (P.K > 0) {
            const int bulk = P.usecutoff ? mbulk : 0;
            for (int qq = dp[u]; qq < dp[u + 1]; qq++) B.bflag[((int64_t)f * P.kb + di[qq]) * P.Mv + v] = bulk;
            for (int qq = dp[v]; qq < dp[v + 1]; qq++) B.bflag[((int64_t)f * P.kb + di[qq]) * P.Mv + u] = bulk;
        }
    }

    __device__ void rdf_hit(int bin) const
    {
        if (bin < P.nbins) atomicAdd(&sh[P.nbins + bin], 1u);
    }

    // Self mode, FP32 stage of the unordered pair {u, v}: both orientations
    // share the distance, the atom pair and the bins of the minimum (the FP64
    // distance is symmetric: the minimum image of -delta is the negated one,
    // rounding included), so md and default groups get +2 in one update; the
    // reference-atom parts belong to their orientation (a reference distance
    // is never below the minimum, so its count needs no test of the main
    // part).  Parts the FP32 interval cannot decide are returned as a mask
    // for the deferred FP64 stage: 1 the minimum, 2 the reference atom of v,
    // 4 the reference atom of u, 8 the whole pair exhaustively (an argmin
    // that FP32 cannot separate).  d0, d1 carry the pair and its argmins.
    __device__ int emit_self(int u, int v, const Best& bm, const Best& brv, const Best& bru, int4& d0, int4& d1) const
    {
        d0 = make_int4(u, v, bm.bi, bm.ba);
        d1 = make_int4(brv.bi, bru.bi, 0, 0);
        const bool amb = (bm.best <= gh.cut2_hi && !(bm.second - bm.best > gh.amb)) ||
                         (brv.best <= gh.cut2_hi && !(brv.second - brv.best > gh.amb)) ||
                         (bru.best <= gh.cut2_hi && !(bru.second - bru.best > gh.amb));
        if (amb) return 8;
        int bm_bin, bm_bulk, rv_bin, ru_bin, dummy;
        const int dmain = bm.best <= gh.cut2_hi ? decide32(P, gh.guard32, bm.best, bm_bin, bm_bulk) : 1;
        const int drv = brv.best <= gh.cut2_hi ? decide32(P, gh.guard32, brv.best, rv_bin, dummy) : 1;
        const int dru = bru.best <= gh.cut2_hi ? decide32(P, gh.guard32, bru.best, ru_bin, dummy) : 1;
        if (dmain == 2) self_main(u, v, bm.bi, bm.ba, bm_bin, bm_bulk);
        if (drv == 2) rdf_hit(rv_bin);
        if (dru == 2) rdf_hit(ru_bin);
        return (dmain == 0 ? 1 : 0) | (drv == 0 ? 2 : 0) | (dru == 0 ? 4 : 0);
    }

    // Self mode, deferred FP64 stage of one pair (mask and entry from
    // emit_self).  The three distances are formed together so their loads
    // are in flight at once; parts outside the mask are discarded.
    __device__ void self_deferred(int4 d0, int4 d1, int mask) const
    {
        const int u = d0.x, v = d0.y;
        if (mask & 8) {
            self_exhaustive(u, v);
            return;
        }
        const int64_t ub = (int64_t)u * P.nu, vb = (int64_t)v * P.nv;
        double a0, a1, a2, b0, b1, b2, c0, c1, c2, e0, e1, e2, p0, p1, p2, q0, q1, q2;
        load3(xuf, ub + max(d0.z, 0), a0, a1, a2);
        load3(xvf, vb + max(d0.w, 0), b0, b1, b2);
        load3(xuf, ub + max(d1.x, 0), c0, c1, c2);
        load3(xvf, vb + P.iref, e0, e1, e2);
        load3(xuf, (int64_t)v * P.nu + max(d1.y, 0), p0, p1, p2);
        load3(xvf, (int64_t)u * P.nv + P.iref, q0, q1, q2);
        const double dm = cm::canon_d2(g.cell, a0, a1, a2, b0, b1, b2);
        const double dr = cm::canon_d2(g.cell, c0, c1, c2, e0, e1, e2);   // atoms of u vs the reference atom of v
        const double du = cm::canon_d2(g.cell, p0, p1, p2, q0, q1, q2);   // atoms of v vs the reference atom of u
        if ((mask & 1) && dm <= P.cut2) {
            const double d = sqrt(dm);
            self_main(u, v, d0.z, d0.w, (int)cm::setbin(d, P.binstep) - 1, d > P.dbulk ? 1 : 0);
        }
        if ((mask & 2) && dr <= P.cut2) rdf_hit((int)cm::setbin(sqrt(dr), P.binstep) - 1);
        if ((mask & 4) && du <= P.cut2) rdf_hit((int)cm::setbin(sqrt(du), P.binstep) - 1);
    }

    // both orientations of a self pair from one set of FP64 distances: the
    // records of (u, v) and (v, u) with their own tie order (update_md,
    // src/minimum_distances.jl:30-39) and reference atoms
    __device__ void self_exhaustive(int u, int v) const
    {
        atomicAdd(&B.stats[0], 2ull);
        Rec r1 = rec_zero(), r2 = rec_zero();
        const int64_t ub = (int64_t)u * P.nu, vb = (int64_t)v * P.nv;
        const int n = NU > 0 ? NU : P.nu;
        auto offer = [&](int i, int a, double x0, double x1, double x2, double y0, double y1, double y2) {
            const double d2 = cm::canon_d2(g.cell, x0, x1, x2, y0, y1, y2);
            if (d2 <= P.cut2) {
                const double d = sqrt(d2);
                rec_offer(r1, i, vb + a, d, a == P.iref);
                rec_offer(r2, a, ub + i, d, i == P.iref);
            }
        };
#pragma unroll 1
        for (int a = 0; a < n; a++) {
            double y0, y1, y2;
            load3(xvf, vb + a, y0, y1, y2);
            if constexpr (NU > 0 && NU <= 4) {
#pragma unroll
                for (int i = 0; i < NU; i++) {
                    double x0, x1, x2;
                    load3(xuf, ub + i, x0, x1, x2);
                    offer(i, a, x0, x1, x2, y0, y1, y2);
                }
            } else {
#pragma unroll 1
                for (int i = 0; i < n; i++) {
                    double x0, x1, x2;
                    load3(xuf, ub + i, x0, x1, x2);
                    offer(i, a, x0, x1, x2, y0, y1, y2);
                }
            }
        }
        rec_complete(P, r1);
        rec_complete(P, r2);
        hist_record_mols(P, sh, shg, B.cnt_gu, B.cnt_gv, r1);
        hist_record_mols(P, sh, shg, B.cnt_gu, B.cnt_gv, r2);
        if (P.K > 0) {
            if (r1.within)
                for (int qq = dp[u]; qq < dp[u + 1]; qq++)
                    B.bflag[((int64_t)f * P.kb + di[qq]) * P.Mv + v] = inbulk(P, r1) ? 1 : 0;
            if (r2.within)
                for (int qq = dp[v]; qq < dp[v + 1]; qq++)
                    B.bflag[((int64_t)f * P.kb + di[qq]) * P.Mv + u] = inbulk(P, r2) ? 1 : 0;
        }
    }

    // one queued pair {u, v, cell of u, shift code or -1}: every shift of the
    // pair, both orientations in self mode
    __device__ int eval_pair(int4 e, int4& d0, int4& d1) const
    {
        const int u = e.x, v = e.y, cu = e.z;
        const int nc0 = gh.nc[0], nc1 = gh.nc[1], nc2 = gh.nc[2];
        int mxs, mys, mzs;   // shift sets (bit m + 1)
        if (e.w >= 0) {
            mxs = 1 << (e.w % 3);
            mys = 1 << ((e.w / 3) % 3);
            mzs = 1 << (e.w / 9);
        } else {
            const int vc = vcellf[v];
            mxs = shifts_onto(vc % nc0, cu % nc0, nc0);
            mys = shifts_onto((vc / nc0) % nc1, (cu / nc0) % nc1, nc1);
            mzs = shifts_onto(vc / (nc0 * nc1), cu / (nc0 * nc1), nc2);
        }
        const int kx = __popc(mxs), ky = __popc(mys), kz = __popc(mzs);
        const float4* xs = solw + (int64_t)u * P.nu;
        const float4* yw = vwf + (int64_t)v * P.nv;
        const float4 cv = vcen[v];
        const float4 cuc = B.ucen[(int64_t)f * P.Mu + u];
        Best bm, brv, bru;
        bm.init(); brv.init(); bru.init();
        if constexpr (NU > 0) {
            // self with NU-atom molecules: both molecules in registers, all
            // loads in flight together
            float4 X[NU], Y[NU];
#pragma unroll
            for (int i = 0; i < NU; i++) X[i] = xs[i];
#pragma unroll
            for (int a = 0; a < NU; a++) Y[a] = yw[a];
            if constexpr (NU <= 4) {
                auto atoms = [&](float tx, float ty, float tz) {
#pragma unroll
                    for (int a = 0; a < NU; a++) {
                        const float yx = Y[a].x - tx, yy = Y[a].y - ty, yz = Y[a].z - tz;
#pragma unroll
                        for (int i = 0; i < NU; i++) {
                            const float ex = yx - X[i].x, ey = yy - X[i].y, ez = yz - X[i].z;
                            const float d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
                            bm.offer(d2, i, a);
                            if (a == P.iref) brv.offer(d2, i, a);
                            if (i == P.iref) bru.offer(d2, a, i);
                        }
                    }
                };
                if (e.w >= 0) {
                    // one shift, whose centre test the walk has already passed
                    // with the same arithmetic (small molecules: a second
                    // inlined copy of the atom loop is cheap)
                    float tx, ty, tz;
                    shift_of(gh, e.w % 3 - 1, (e.w / 3) % 3 - 1, e.w / 9 - 1, tx, ty, tz);
                    atoms(tx, ty, tz);
                } else {
                    for (int iz = 0; iz < kz; iz++)
                    for (int iy = 0; iy < ky; iy++)
                    for (int ix = 0; ix < kx; ix++) {
                        float tx, ty, tz;
                        shift_of(gh, nth_shift(mxs, ix), nth_shift(mys, iy), nth_shift(mzs, iz), tx, ty, tz);
                        const float dx = cv.x - (cuc.x + tx), dy = cv.y - (cuc.y + ty), dz = cv.z - (cuc.z + tz);
                        if (dx * dx + dy * dy + dz * dz > gh.rpre2) continue;
                        atoms(tx, ty, tz);
                    }
                }
            } else {
                for (int iz = 0; iz < kz; iz++)
                for (int iy = 0; iy < ky; iy++)
                for (int ix = 0; ix < kx; ix++) {
                    float tx, ty, tz;
                    shift_of(gh, nth_shift(mxs, ix), nth_shift(mys, iy), nth_shift(mzs, iz), tx, ty, tz);
                    const float dx = cv.x - (cuc.x + tx), dy = cv.y - (cuc.y + ty), dz = cv.z - (cuc.z + tz);
                    if (dx * dx + dy * dy + dz * dz > gh.rpre2) continue;
#pragma unroll
                    for (int a = 0; a < NU; a++) {
                        const float yx = Y[a].x - tx, yy = Y[a].y - ty, yz = Y[a].z - tz;
#pragma unroll
                        for (int i = 0; i < NU; i++) {
                            const float ex = yx - X[i].x, ey = yy - X[i].y, ez = yz - X[i].z;
                            const float d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
                            bm.offer(d2, i, a);
                            if (a == P.iref) brv.offer(d2, i, a);
                            if (i == P.iref) bru.offer(d2, a, i);
                        }
                    }
                }
            }
        } else {
            for (int iz = 0; iz < kz; iz++)
            for (int iy = 0; iy < ky; iy++)
            for (int ix = 0; ix < kx; ix++) {
                float tx, ty, tz;
                shift_of(gh, nth_shift(mxs, ix), nth_shift(mys, iy), nth_shift(mzs, iz), tx, ty, tz);
                const float dx = cv.x - (cuc.x + tx), dy = cv.y - (cuc.y + ty), dz = cv.z - (cuc.z + tz);
                if (dx * dx + dy * dy + dz * dz > gh.rpre2) continue;
                for (int a = 0; a < P.nv; a++) {
                    const float4 y = yw[a];
                    const float yx = y.x - tx, yy = y.y - ty, yz = y.z - tz;
                    for (int i = 0; i < P.nu; i++) {
                        const float4 s4 = xs[i];
                        const float ex = yx - s4.x, ey = yy - s4.y, ez = yz - s4.z;
                        const float d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
                        bm.offer(d2, i, a);
                        if (a == P.iref) brv.offer(d2, i, a);
                        if (P.self_ && i == P.iref) bru.offer(d2, a, i);
                    }
                }
            }
        }
        if (P.self_) return emit_self(u, v, bm, brv, bru, d0, d1);
        emit(u, v, bm, brv);
        return 0;
    }
};

// Real pass of the many-solute-molecule family.  blockIdx.y is the frame;
// each lane owns one solvent molecule v (self: in cell order).  The lanes walk
// their distinct neighbour cells together (trip count = the wavefront's
// largest cell, four candidates per step with their loads in flight), the
// molecule pairs with a centre within reach under some lattice shift are
// appended to an LDS queue (ballot offsets), and the queue is evaluated 64
// pairs at a time, one pair per lane, so the nu x nv atom loops (over every
// shift that maps the two cells onto each other) run without divergence.
// Self-MDDF evaluates each unordered pair once (u < v) and emits the records
// of both orientations (same distance and argmin pair; reference atoms and tie
// order per orientation).
template <class TU, class TV, int NU>
__global__ void __launch_bounds__(BLOCK, 3) k_eval_mols_real(KP P, Bufs B, const TU* xu, const TV* xv, int gx, int S)
{
    // XCD-aware order: workgroups go round-robin to the 8 XCDs, each with its
    // own L2.  XCD k takes the k-th eighth of the (frame, block) items in
    // frame-major order, so each XCD works through its frames one at a time
    // and its L2 holds the molecules of one frame instead of a slice of
    // every frame of the batch.  gx blocks cover one frame.
    const int items = gx * P.F;
    const int per = (items + 7) >> 3;
    const int item = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (item >= items) return;   // uniform per block, before any barrier
    const int f = item / gx, bx = item - (item / gx) * gx;
    extern __shared__ unsigned sh[];
    const int gwords = P.glds ? P.nbins * (P.gnu + (P.self_ ? 0 : P.gnv)) : 0;
    hist_clear(sh, 2 * P.nbins + gwords);
    unsigned* shg = P.glds ? sh + 2 * P.nbins : nullptr;
    int4* q = (int4*)(sh + ((2 * P.nbins + gwords + 3) & ~3)) + (threadIdx.x >> 6) * QCAP;
    // self: deferred FP64 entries (two int4 each) after the pair queues
    int4* dq = (int4*)(sh + ((2 * P.nbins + gwords + 3) & ~3)) + (BLOCK / 64) * QCAP + (threadIdx.x >> 6) * 2 * DCAP;
    int dqn = 0;
    const int lane = threadIdx.x & 63;
    const PairCtx<TU, TV, NU> ctx(P, B, f, xu, xv, sh, shg);
    const GeoHot& gh = ctx.gh;
    const float4* srt = B.sorted + (int64_t)f * B.nitems;
    const int* st = B.start + (int64_t)f * (P.cap + 1);
    const int nc[3] = {gh.nc[0], gh.nc[1], gh.nc[2]};
    const int nd[3] = {min(nc[0], 3), min(nc[1], 3), min(nc[2], 3)};
    const bool simple = nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3;
    constexpr int CU = 8;   // candidates per lane per step
    const int stride = gx * blockDim.x;
    const unsigned long long lt = (1ull << lane) - 1ull;
    // S lanes per solvent molecule (small systems): lane sub of a molecule
    // walks the offsets sub, sub + S, ... of the neighbour walk
    for (int base = bx * blockDim.x + (threadIdx.x & ~63); base < P.Mv * S; base += stride) {
        const int vl = base + lane;
        const int q0 = vl / S, sub = vl - (vl / S) * S;
        const bool act = q0 < P.Mv;
        const int v = act ? (P.self_ ? __float_as_int(srt[q0].w) : q0) : 0;
        const float4 cv = ctx.vcen[v];
        const int vcell = ctx.vcellf[v];
        const int cc[3] = {vcell % nc[0], (vcell / nc[0]) % nc[1], vcell / (nc[0] * nc[1])};
        const int ncells = nd[0] * nd[1] * nd[2];
        // The walk over the neighbour cells produces queue entries until 64
        // are waiting, the wavefront evaluates them, and the walk resumes where
        // it stopped: the pair evaluation has one call site (one inlined copy).
        int qn = 0;
        // Self with three or more cells per dimension: every offset maps to a
        // distinct cell, so each unordered pair of molecules in different
        // cells is met exactly once from the half shell of offsets after
        // (0, 0, 0) (linear index 13 of the 27; z, then y, then x), and pairs
        // inside one cell once with u < v.  Otherwise all offsets with u < v.
        const int o_first = (P.self_ && simple) ? 13 : 0;
        const int nsteps = (ncells - o_first + S - 1) / S;   // walk steps, the same for every lane
        int ks = 0, o = o_first + sub, kpos = 0, kmax = -1;
        int cell = 0, cs = 0, cnt = 0, code = -1, kx = 1, ky = 1, kz = 1;
        int sxs = 0, sys_ = 0, szs = 0;   // shift sets of the current cell (small grids)
        float t1x = 0.f, t1y = 0.f, t1z = 0.f;   // regular grids: own centre minus the cell's shift
        for (;;) {
            while (qn < 64 && ks < nsteps) {
                if (kmax < 0) {
                    // enter neighbour cell o: offsets -1..1, or 0..nc-1 when nc < 3
                    // (a lane whose offsets are used up walks nothing)
                    const bool ov = o < ncells;
                    const int oc = ov ? o : o_first;
                    const int ox = oc % nd[0], oy = (oc / nd[0]) % nd[1], oz = oc / (nd[0] * nd[1]);
                    int m0;
                    const int cx = nb_cell(cc[0], nc[0] >= 3 ? ox - 1 : ox, nc[0], m0);
                    const int cy = nb_cell(cc[1], nc[1] >= 3 ? oy - 1 : oy, nc[1], m0);
                    const int cz = nb_cell(cc[2], nc[2] >= 3 ? oz - 1 : oz, nc[2], m0);
                    cell = (cz * nc[1] + cy) * nc[0] + cx;
                    if (simple) {
                        // three or more cells per dimension: one shift
                        int mx, my, mz;
                        nb_cell(cc[0], ox - 1, nc[0], mx);
                        nb_cell(cc[1], oy - 1, nc[1], my);
                        nb_cell(cc[2], oz - 1, nc[2], mz);
                        code = (mx + 1) + 3 * (my + 1) + 9 * (mz + 1);
                        shift_of(gh, mx, my, mz, t1x, t1y, t1z);
                        // the centre test below is a pre-filter with a margin
                        // (rpre), so its rounding does not matter: the shift is
                        // taken off the lane's own centre once per cell
                        t1x = cv.x - t1x;
                        t1y = cv.y - t1y;
                        t1z = cv.z - t1z;
                    } else {
                        sxs = shifts_onto(cc[0], cx, nc[0]);
                        sys_ = shifts_onto(cc[1], cy, nc[1]);
                        szs = shifts_onto(cc[2], cz, nc[2]);
                        kx = __popc(sxs);
                        ky = __popc(sys_);
                        kz = __popc(szs);
                    }
                    cs = st[cell];
                    cnt = (act && ov) ? st[cell + 1] - cs : 0;
                    kmax = wave_max_i(cnt);
                    kpos = 0;
                }
                if (kpos >= kmax) {
                    ks++;
                    o += S;
                    kmax = -1;
                    continue;
                }
                float4 cu4[CU];
#pragma unroll
                for (int w = 0; w < CU; w++) cu4[w] = srt[min(cs + min(kpos + w, max(cnt - 1, 0)), P.Mu - 1)];
                int w = 0;
#pragma unroll
                for (int ww = 0; ww < CU; ww++) {
                    if (w != ww || kpos + ww >= kmax || qn >= 64) break;
                    const int u = __float_as_int(cu4[ww].w);
                    bool pass = kpos + ww < cnt && (P.self_ ? ((simple && o > 13) || u < v) : true);
                    if (simple) {
                        const float dx = t1x - cu4[ww].x, dy = t1y - cu4[ww].y, dz = t1z - cu4[ww].z;
                        pass = pass && dx * dx + dy * dy + dz * dz <= gh.rpre2;
                    } else if (pass) {
                        bool nearby = false;
                        for (int iz = 0; iz < kz; iz++)
                        for (int iy = 0; iy < ky; iy++)
                        for (int ix = 0; ix < kx; ix++) {
                            float tx, ty, tz;
                            shift_of(gh, nth_shift(sxs, ix), nth_shift(sys_, iy), nth_shift(szs, iz), tx, ty, tz);
                            const float dx = cv.x - (cu4[ww].x + tx), dy = cv.y - (cu4[ww].y + ty),
                                        dz = cv.z - (cu4[ww].z + tz);
                            nearby |= dx * dx + dy * dy + dz * dz <= gh.rpre2;
                        }
                        pass = nearby;
                    }
                    const unsigned long long m = __ballot(pass);
                    if (pass) q[qn + __popcll(m & lt)] = make_int4(u, v, cell, code);
                    qn += __popcll(m);
                    w = ww + 1;
                }
                kpos += w;
            }
            if (qn == 0) break;
            const int ne = min(qn, 64);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const bool has = lane < ne;
            const int4 e = q[has ? qn - ne + lane : 0];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            qn -= ne;
            int4 d0 = make_int4(0, 0, 0, 0), d1 = d0;
            int dm = 0;
            if (has) dm = ctx.eval_pair(e, d0, d1);
            if (P.self_) {
                // parts FP32 could not decide wait in the deferred queue and
                // are resolved 64 at a time, so the FP64 work runs on full
                // wavefronts
                const unsigned long long m = __ballot(dm != 0);
                if (dm != 0) {
                    const int k = dqn + __popcll(m & lt);
                    dq[2 * k] = d0;
                    dq[2 * k + 1] = make_int4(d1.x, d1.y, dm, 0);
                }
                dqn += __popcll(m);
                if (dqn >= 64) {
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    const int k = dqn - 64 + lane;
                    const int4 x0 = dq[2 * k], x1 = dq[2 * k + 1];
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    dqn -= 64;
                    ctx.self_deferred(x0, x1, x1.z);
                }
            }
        }
    }
    if (dqn > 0) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane < dqn) {
            const int4 x0 = dq[2 * lane], x1 = dq[2 * lane + 1];
            ctx.self_deferred(x0, x1, x1.z);
        }
    }
    hist_flush(sh, P.nbins, B.cnt_md, B.cnt_rdf);
    if (shg) groups_flush(P, shg, B.cnt_gu, B.cnt_gv);
}

// Whether placed slot v (centre cv) can have an atom pair with the drawn
// solute molecule (centre cu) within the cutoff under some lattice shift:
// the centre pre-check of the random passes of the molecule-pair family.
// k_place and k_eval_mols_rand evaluate it with the same FP32 operations on
// the same values, so a slot skipped by the placement is one the evaluation
// skips.
__device__ __forceinline__ bool mol_pair_near(const Geo& g, float4 cv, float4 cu)
{
    bool any = false;
    for (int mz = -1; mz <= 1; mz++)
    for (int my = -1; my <= 1; my++)
    for (int mx = -1; mx <= 1; mx++) {
        float fx = (float)mx, fy = (float)my, fz = (float)mz;
        float tx = g.H32[0] * fx + g.H32[3] * fy + g.H32[6] * fz;
        float ty = g.H32[1] * fx + g.H32[4] * fy + g.H32[7] * fz;
        float tz = g.H32[2] * fx + g.H32[5] * fy + g.H32[8] * fz;
        float dx = cv.x - (cu.x + tx), dy = cv.y - (cu.y + ty), dz = cv.z - (cu.z + tz);
        any |= !(dx * dx + dy * dy + dz * dz > g.rpre2);
    }
    return any;
}

template <class TU>
__global__ void __launch_bounds__(BLOCK) k_eval_mols_rand(KP P, Bufs B, const TU* xu)
{
    extern __shared__ unsigned sh[];
    const int gwords = P.glds ? P.nbins * (P.gnu + (P.self_ ? 0 : P.gnv)) : 0;
    hist_clear(sh, 2 * P.nbins + gwords);
    unsigned* shg = P.glds ? sh + 2 * P.nbins : nullptr;
    int64_t total = (int64_t)P.F * P.K * P.Mv;
    for (int64_t it = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; it < total; it += (int64_t)gridDim.x * blockDim.x) {
        int64_t fk = it / P.Mv;
        int v = (int)(it - fk * P.Mv);
        int f = (int)(fk / P.K);
        int u = B.uk[fk];
        if (P.self_ && u == v) continue;
        const Geo& g = B.geo[f];
        const float4* yw = B.vw_r + fk * P.Nv + (int64_t)v * P.nv;
        float4 cv = B.vcen_r[fk * P.Mv + v];
        float4 cu = B.ucen[(int64_t)f * P.Mu + u];
        const float4* xs = B.solw + (int64_t)f * P.Nu + (int64_t)u * P.nu;
        Best bm, br;
        bm.init(); br.init();
        // the slots k_place skipped fail this test (the same one)
        if (!mol_pair_near(g, cv, cu)) continue;
        for (int mz = -1; mz <= 1; mz++)
        for (int my = -1; my <= 1; my++)
        for (int mx = -1; mx <= 1; mx++) {
            float fx = (float)mx, fy = (float)my, fz = (float)mz;
            float tx = g.H32[0] * fx + g.H32[3] * fy + g.H32[6] * fz;
            float ty = g.H32[1] * fx + g.H32[4] * fy + g.H32[7] * fz;
            float tz = g.H32[2] * fx + g.H32[5] * fy + g.H32[8] * fz;
            float dx = cv.x - (cu.x + tx), dy = cv.y - (cu.y + ty), dz = cv.z - (cu.z + tz);
            if (dx * dx + dy * dy + dz * dz > g.rpre2) continue;
            pairs_mol(P, xs, yw, tx, ty, tz, bm, br);
        }
        const TU* xuf = xu + (int64_t)f * P.Nu * 3;
        int64_t jbase = (int64_t)v * P.nv;
        auto ysrc = [&](int a, double& y0, double& y1, double& y2) {
            const double* pp = B.P + (fk * P.Nv + jbase + a) * 3; y0 = pp[0]; y1 = pp[1]; y2 = pp[2];
        };
        Rec r = resolve_mols(P, g, xuf, u, v, ysrc, bm, br, &B.stats[0]);
        hist_record_mols(P, sh, shg, B.cnt_gu_r, B.cnt_gv_r, r);
    }
    hist_flush(sh, P.nbins, B.cnt_md_r, B.cnt_rdf_r);
    if (shg) groups_flush(P, shg, B.cnt_gu_r, B.cnt_gv_r);
}

// ---------------------------------------------------------------------------
// bulk lists and random placement
// ---------------------------------------------------------------------------
__device__ __forceinline__ void b_bulk_init(KP P, Bufs B, int bx_, int nbx_, int by_, int nby_)
{
    int64_t total = (int64_t)P.F * P.kb * P.Mv;
    for (int64_t it = bx_ * (int64_t)blockDim.x + threadIdx.x; it < total; it += (int64_t)nbx_ * blockDim.x) {
        int64_t fk = it / P.Mv;
        int v = (int)(it - fk * P.Mv);
        unsigned char fl = P.usecutoff ? 0 : 1;
        if (P.self_ && B.uk[fk] == v) fl = 0;   // the bulk pool excludes isolute, src/mddf.jl:410
        B.bflag[it] = fl;
    }
}

__global__ void k_bulk_init(KP P, Bufs B)
{
    b_bulk_init(P, B, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// ordered compaction of the bulk flags, one block per (frame, list)
__device__ __forceinline__ void b_bulk_compact(KP P, Bufs B, int bx_, int nbx_, int by_, int nby_)
{
    __shared__ int wsum[16];
    __shared__ int carry;
    int64_t fk = bx_;
    const unsigned char* fl = B.bflag + fk * P.Mv;
    int* out = B.blist + fk * P.Mv;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // four consecutive flags per thread, the counts scanned over the
    // wavefront (DPP) and across the block's wavefronts
    for (int base = 0; base < P.Mv; base += 4 * blockDim.x) {
        const int v0 = base + 4 * threadIdx.x;
        int fl4[4];
#pragma unroll
        for (int u = 0; u < 4; u++) fl4[u] = v0 + u < P.Mv ? fl[v0 + u] : 0;
        const int cnt = fl4[0] + fl4[1] + fl4[2] + fl4[3];
        const int incl = wave_incl_scan(cnt);
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w = 0; w < nw; w++) { if (w < wid) woff += wsum[w]; tot += wsum[w]; }
        int o = carry + woff + incl - cnt;
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (fl4[u]) out[o++] = v0 + u;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) B.nbulk[fk] = carry;
}

__global__ void __launch_bounds__(1024) k_bulk_compact(KP P, Bufs B)
{
    b_bulk_compact(P, B, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// The atoms of one source molecule (xyz interleaved) and the placed atoms
// (FP64 for the exact resolve, FP32 for the evaluation), through global
// pointers (also where the pointers were read from a job table).
template <class TV>
struct MolSrc {
    gptr<const TV> p;
    __device__ MolSrc(const TV* q) : p((gptr<const TV>)q) {}
    __device__ double operator()(int a, int k) const { return (double)p[3 * a + k]; }
};

struct MolDst {
    gptr<double> p;
    gptr<f4v> w;
    __device__ MolDst(double* q, float4* v) : p((gptr<double>)q), w((gptr<f4v>)v) {}
    __device__ void operator()(int a, double x, double y, double z) const
    {
        p[3 * a] = x; p[3 * a + 1] = y; p[3 * a + 2] = z;
        w[a] = f4v{(float)x, (float)y, (float)z, 0.f};
    }
    // the first atom's fourth FP32 component
    __device__ void set_w0(float x) const { ((gptr<float>)w)[3] = x; }
};

// randomize_solvent!, src/mddf.jl:65-88, for every slot of every random pass
template <class TV>
__global__ void k_place(KP P, Bufs B, const TV* xv)
{
    // grid: x over the slots, y over the (frame, pass) pairs
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.Mv) return;
    for (int fk = blockIdx.y; fk < P.F * P.K; fk += gridDim.y) {
        const int f = fk / P.K;
        const int k = fk - f * P.K;
        const Geo& g = B.geo[f];
        cm::u32x4 r = cm::philox4x32_10({(uint32_t)v, (uint32_t)k, g.frame, cm::TAG_SLOT}, P.key0, P.key1);
        double uc[3] = {cm::u01(r.y), cm::u01(r.z), cm::u01(r.w)};
        // the new centre H*uc (inside the cell)
        const double* H = g.cell.H;
        double c0 = H[0] * uc[0] + H[3] * uc[1] + H[6] * uc[2];
        double c1 = H[1] * uc[0] + H[4] * uc[1] + H[7] * uc[2];
        double c2 = H[2] * uc[0] + H[5] * uc[1] + H[8] * uc[2];
        const float4 cv = mk4((float)c0, (float)c1, (float)c2, 0.f);
        B.vcen_r[(int64_t)fk * P.Mv + v] = cv;
        // only slots the evaluation can reach are placed (all of them when
        // the placements are kept for verification)
        if (!P.place_all) {
            const int u = B.uk[fk];
            if (P.self_ && u == v) continue;
            if (!mol_pair_near(g, cv, B.ucen[(int64_t)f * P.Mu + u])) continue;
        }
        cm::u32x4 r2 = cm::philox4x32_10({(uint32_t)v, (uint32_t)k, g.frame, cm::TAG_SLOT + 1u}, P.key0, P.key1);
        int64_t lk = (int64_t)f * P.kb + (P.kb == 1 ? 0 : k);
        int nb = B.nbulk[lk];
        int64_t jmol = nb > 0 ? (int64_t)B.blist[lk * P.Mv + cm::draw_index(r.x, nb)] : cm::draw_index(r.x, P.Mv);
        MolSrc<TV> src(xv + (int64_t)f * P.Nv * 3 + jmol * P.nv * 3);
        MolDst dst(B.P + (fk * P.Nv + (int64_t)v * P.nv) * 3, B.vw_r + fk * P.Nv + (int64_t)v * P.nv);
        double ua[3] = {cm::u01(r2.x), cm::u01(r2.y), cm::u01(r2.z)};
        cm::random_move(src, dst, P.nv, P.iref, g.cell, uc, ua);
    }
}

// ---------------------------------------------------------------------------
// tiles family: one solute molecule, solvent molecules of up to 64 atoms
// ---------------------------------------------------------------------------
constexpr int TCAP_DEFAULT = 640;   // staged solute atoms per round (10 KiB of LDS)
constexpr int RB_PIECES = 8;        // read-back pieces of cm_job_finish
constexpr int TCAP_MOL = 512;       // the same with the squared norms (20 B each: 10 KiB)
constexpr int CBCAP = 1024;         // chunk-table entries kept in LDS
// The last DYN_PCT per cent of an evaluation launch's chunks are handed out
// on request, in launches of at least DYN_MIN chunks per workgroup, in about
// DYN_GRABS requests per workgroup (consecutive chunks per request).  On the
// 200-frame globule job and 64 uneven systems (random-pass launch time):
// 35% in requests of one chunk -9%, of two (8 requests) -17%, of four -12%;
// 25% -11%, 45% (two per request) -17%, 55% and more in single chunks slower
// than a fixed stride alone.
constexpr int DYN_PCT = 35;
constexpr int DYN_MIN = 2;
constexpr int DYN_GRABS = 8;
constexpr int SU = 4;               // staging: atoms per lane per step
constexpr int SEG = 8;              // staging: atoms of a row per lane (segment)

__device__ __forceinline__ int wrap_index(int i, int n, int& m)
{
    int q = i >= 0 ? i / n : -((-i + n - 1) / n);
    m = q;
    return i - q * n;
}

// Morton code of a fine cell: chunks of consecutive molecules in this order
// stay compact in space
// bits 0..9 of x moved to bits 0, 3, 6, ..., 27
__device__ __forceinline__ unsigned spread3(unsigned x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Morton code of the fine cell (c0, c1, c2), each below 2^kbits
__device__ __forceinline__ int morton_key3(const Geo& g, int c0, int c1, int c2)
{
    if (g.kbits <= 10) return (int)(spread3(c0) | (spread3(c1) << 1) | (spread3(c2) << 2));
    unsigned k = 0;
    for (int b = 0; b < g.kbits; b++) {
        k |= (((unsigned)c0 >> b) & 1u) << (3 * b);
        k |= (((unsigned)c1 >> b) & 1u) << (3 * b + 1);
        k |= (((unsigned)c2 >> b) & 1u) << (3 * b + 2);
    }
    return (int)k;
}

// Fine-grid index ranges covering the box [lo - r, hi + r] (fractional
// coordinates of its 8 corners, with a small outward margin).
__device__ __forceinline__ void cell_range(const Geo& g, float lx, float ly, float lz, float hx, float hy, float hz,
                                           float r, int ilo[3], int ihi[3])
{
    const double* Hi = g.cell.Hi;
    double fmn[3] = {INFINITY, INFINITY, INFINITY}, fmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (g.ortho) {
        // diagonal cell: the corners' fractional coordinates are the two
        // products per dimension (the off-diagonal terms are exact zeros, so
        // the values equal the general form's)
        const double lo3[3] = {(double)lx - r, (double)ly - r, (double)lz - r};
        const double hi3[3] = {(double)hx + r, (double)hy + r, (double)hz + r};
        for (int k = 0; k < 3; k++) {
            const double a = Hi[4 * k] * lo3[k], b = Hi[4 * k] * hi3[k];
            ilo[k] = (int)floor(fmin(a, b) * g.nc[k] - 1e-6);
            ihi[k] = (int)floor(fmax(a, b) * g.nc[k] + 1e-6);
        }
        return;
    }
    for (int c = 0; c < 8; c++) {
        double x = (c & 1) ? (double)hx + r : (double)lx - r;
        double y = (c & 2) ? (double)hy + r : (double)ly - r;
        double z = (c & 4) ? (double)hz + r : (double)lz - r;
        double f[3] = {Hi[0] * x + Hi[3] * y + Hi[6] * z, Hi[1] * x + Hi[4] * y + Hi[7] * z,
                       Hi[2] * x + Hi[5] * y + Hi[8] * z};
        for (int k = 0; k < 3; k++) { fmn[k] = fmin(fmn[k], f[k]); fmx[k] = fmax(fmx[k], f[k]); }
    }
    for (int k = 0; k < 3; k++) {
        ilo[k] = (int)floor(fmn[k] * g.nc[k] - 1e-6);
        ihi[k] = (int)floor(fmx[k] * g.nc[k] + 1e-6);
    }
}


// cells near the solute: every cell within rnear of a solute atom (index
// box).  A block takes 256 consecutive atoms of one frame (neighbours along
// the chain, whose cell ranges overlap): the marks go into a bitmap of the
// frame's cells in LDS first, and the block writes each marked cell once.
constexpr int NEAR_WORDS = 8192;   // LDS bitmap words (frames of up to 262,144 cells; larger: direct stores)

__device__ __forceinline__ void b_near_mark(KP P, Bufs B, int bx_, int nbx_, int by_, int nby_)
{
    __shared__ unsigned bits[NEAR_WORDS];
    for (int f = by_; f < P.F; f += nby_) {
        const Geo g = ldc_geo(B.geo + f);
        const int ncell = g.nc[0] * g.nc[1] * g.nc[2];
        const bool lds = ncell <= NEAR_WORDS * 32;
        const int nw = (ncell + 31) >> 5;
        if (lds)
            for (int w = threadIdx.x; w < nw; w += blockDim.x) bits[w] = 0u;
        __syncthreads();
        unsigned char* nr = B.near + (int64_t)f * P.cap;
        // boxes of a diagonal cell: only the cells whose box is within rnear
        // (in the image of index i); otherwise the whole index box
        const float rn2 = g.rnear * g.rnear;
        const float hx = g.H32[0] / g.nc[0], hy = g.H32[4] / g.nc[1], hz = g.H32[8] / g.nc[2];
        auto gap = [](float v, float lo, float hi) { return fmaxf(fmaxf(lo - v, v - hi), 0.f); };
        // cells a..b (0 <= a <= b < nc0) of the row starting at cell rb: a
        // bit range, one LDS atomic per word
        auto mark_run = [&](int rb, int a, int b) {
            const int c0 = rb + a, c1 = rb + b;
            if (lds) {
                for (int w = c0 >> 5; w <= (c1 >> 5); w++) {
                    const int lo = w == (c0 >> 5) ? (c0 & 31) : 0, hi = w == (c1 >> 5) ? (c1 & 31) : 31;
                    const unsigned m = (hi == 31 ? 0xffffffffu : (2u << hi) - 1u) & ~((1u << lo) - 1u);
                    atomicOr(&bits[w], m);
                }
            } else {
                for (int c = c0; c <= c1; c++) nr[c] = 1;
            }
        };
        for (int64_t at = bx_ * (int64_t)blockDim.x + threadIdx.x; at < P.nu; at += (int64_t)nbx_ * blockDim.x) {
            const float4 s = B.tmp4[(int64_t)f * P.nu + at];
            int ilo[3], ihi[3];
            cell_range(g, s.x, s.y, s.z, s.x, s.y, s.z, g.rnear, ilo, ihi);
            for (int iz = ilo[2]; iz <= ihi[2]; iz++) {
                int mz; int cz = wrap_index(iz, g.nc[2], mz);
                const float dz = g.ortho ? gap(s.z, iz * hz, (iz + 1) * hz) : 0.f;
                for (int iy = ilo[1]; iy <= ihi[1]; iy++) {
                    int my; int cy = wrap_index(iy, g.nc[1], my);
                    const float dy = g.ortho ? gap(s.y, iy * hy, (iy + 1) * hy) : 0.f;
                    const float r2 = rn2 - dz * dz - dy * dy;
                    if (r2 < 0.f) continue;
                    // the x cells of the row within rnear (diagonal cells: their
                    // box gap <= sqrt(r2), widened over FP32 rounding, which
                    // only adds cells), as at most two runs of wrapped cells
                    int x0 = ilo[0], x1 = ihi[0];
                    if (g.ortho) {
                        const float rx = sqrtf(r2) * (1.0f + 1e-6f) + 1e-4f;
                        x0 = max(x0, (int)floorf((s.x - rx) / hx));
                        x1 = min(x1, (int)floorf((s.x + rx) / hx));
                    }
                    if (x0 > x1) continue;
                    const int rb = (cz * g.nc[1] + cy) * g.nc[0];
                    if (x1 - x0 + 1 >= g.nc[0]) {
                        mark_run(rb, 0, g.nc[0] - 1);
                    } else {
                        int m0, m1;
                        const int a = wrap_index(x0, g.nc[0], m0), b = wrap_index(x1, g.nc[0], m1);
                        if (a <= b) {
                            mark_run(rb, a, b);
                        } else {
                            mark_run(rb, a, g.nc[0] - 1);
                            mark_run(rb, 0, b);
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (lds)
            for (int w = threadIdx.x; w < nw; w += blockDim.x) {
                unsigned m = bits[w];
                while (m) {
                    const int bit = __ffs(m) - 1;
                    m &= m - 1u;
                    nr[(w << 5) + bit] = 1;
                }
            }
        __syncthreads();
    }
}

__global__ void k_near_mark(KP P, Bufs B)
{
    b_near_mark(P, B, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// Every near cell: the distance from its centre q to the nearest image of a
// solute atom in the cell itself or one of its six face neighbours, and the
// distance from its lower corner to the nearest image in the eight cells
// around that corner; each encoded as 0xffffffff - bits (0, the cleared
// value, when the cells hold no atom).  Each is the distance to one
// particular atom image, so for any point p min_a |p - a| <= |p - q| + that
// distance: each molecule of the cell gets from them an upper bound on its
// atoms' nearest solute distance (molecule_bound).  Sixteen lanes per cell:
// lane j < 7 scans one cell for the centre, lane 8 + t one for the corner.
__device__ __forceinline__ void b_dfield(KP P, Bufs B, int bx_, int nbx_, int by_, int nby_)
{
    const int64_t total = (int64_t)P.F * P.cap;
    const int j = threadIdx.x & 15;
    const int64_t step = ((int64_t)nbx_ * blockDim.x) >> 4;
    // the sixteen lanes of a group share it (blockDim is a multiple of 16)
    for (int64_t it = (bx_ * (int64_t)blockDim.x + threadIdx.x) >> 4; it < total; it += step) {
        const bool nearc = B.near[it] != 0;
        float best = INFINITY;
        if (nearc && j != 7) {
            const int f = (int)(it / P.cap);
            const int cell = (int)(it - (int64_t)f * P.cap);
            const Geo g = ldc_geo(B.geo + f);
            const int cx = cell % g.nc[0], cy = (cell / g.nc[0]) % g.nc[1], cz = cell / (g.nc[0] * g.nc[1]);
            int ax, ay, az;
            float fx, fy, fz;
            if (j < 7) {
                ax = j == 1 ? 1 : (j == 2 ? -1 : 0);
                ay = j == 3 ? 1 : (j == 4 ? -1 : 0);
                az = j == 5 ? 1 : (j == 6 ? -1 : 0);
                fx = (cx + 0.5f) / g.nc[0]; fy = (cy + 0.5f) / g.nc[1]; fz = (cz + 0.5f) / g.nc[2];
            } else {
                const int t = j - 8;
                ax = -(t & 1); ay = -((t >> 1) & 1); az = -((t >> 2) & 1);
                fx = (float)cx / g.nc[0]; fy = (float)cy / g.nc[1]; fz = (float)cz / g.nc[2];
            }
            int mx, my, mz;
            const int ix = wrap_index(cx + ax, g.nc[0], mx), iy = wrap_index(cy + ay, g.nc[1], my),
                      iz = wrap_index(cz + az, g.nc[2], mz);
            float tx, ty, tz;
            shift_of(g, mx, my, mz, tx, ty, tz);
            // the reference point minus the image shift, against the wrapped atoms
            const float rx = g.H32[0] * fx + g.H32[3] * fy + g.H32[6] * fz - tx;
            const float ry = g.H32[1] * fx + g.H32[4] * fy + g.H32[7] * fz - ty;
            const float rz = g.H32[2] * fx + g.H32[5] * fy + g.H32[8] * fz - tz;
            const int* st = B.start + (int64_t)f * (P.cap + 1);
            const float4* srt = B.sorted + (int64_t)f * B.nitems;
            const int c2 = (iz * g.nc[1] + iy) * g.nc[0] + ix;
            const int k0 = st[c2], k1 = st[c2 + 1];
            int k = k0;
            for (; k + 3 < k1; k += 4) {
                float4 a[4];
#pragma unroll
                for (int u = 0; u < 4; u++) a[u] = srt[k + u];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const float ex = a[u].x - rx, ey = a[u].y - ry, ez = a[u].z - rz;
                    best = fminf(best, ex * ex + ey * ey + ez * ez);
                }
            }
            for (; k < k1; k++) {
                const float4 a = srt[k];
                const float ex = a.x - rx, ey = a.y - ry, ez = a.z - rz;
                best = fminf(best, ex * ex + ey * ey + ez * ez);
            }
        }
        // minimum over each eight-lane half (centre: lanes 0-6, corner: 8-15)
        best = fminf(best, __shfl_xor(best, 1, 64));
        best = fminf(best, __shfl_xor(best, 2, 64));
        best = fminf(best, __shfl_xor(best, 4, 64));
        if (nearc && (j & 7) == 0 && best < INFINITY)
            B.dfield[(j ? total : 0) + it] = 0xffffffffu - __float_as_uint(sqrtf(best));
    }
}

__global__ void k_dfield(KP P, Bufs B)
{
    b_dfield(P, B, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// Upper bound (FP32, rounded outward by a wide margin) on the distance from
// any atom of a molecule whose centre has fractional coordinates fr (in the
// near cell `cell`) to its nearest solute atom: the smallest, over the cell's
// centre and its eight corners, of the reference point's atom distance plus
// the distance from the centre to the point, plus the molecule radius; +inf
// when none has a distance.  FP32 throughout: every rounding is far below the
// 1e-5 relative + 0.01 A margin.
struct BoundLoads {
    unsigned e[9];   // the centre's entry, then the corners (a + 2 b + 4 c)
};

// the entries, read by the callers before their stores (which may alias them
// as far as the compiler knows, and would hold the loads back)
__device__ __forceinline__ BoundLoads bound_loads(const KP& P, const Bufs& B, const Geo& g, int f, int cell, int c0,
                                                  int c1, int c2)
{
    // cell = (c2 nc1 + c1) nc0 + c0
    BoundLoads L;
    const gptr<const unsigned> d = (gptr<const unsigned>)(B.dfield + (int64_t)f * P.cap);
    const gptr<const unsigned> dc = d + (int64_t)P.F * P.cap;
    L.e[0] = d[cell];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        int i0 = c0 + (t & 1), i1 = c1 + ((t >> 1) & 1), i2 = c2 + ((t >> 2) & 1);
        if (i0 == g.nc[0]) i0 = 0;
        if (i1 == g.nc[1]) i1 = 0;
        if (i2 == g.nc[2]) i2 = 0;
        L.e[1 + t] = dc[(i2 * g.nc[1] + i1) * g.nc[0] + i0];
    }
    return L;
}

__device__ __forceinline__ float molecule_bound(const Geo& g, const BoundLoads& L, int c0, int c1, int c2,
                                                const double fr[3])
{
    const float* H = g.H32;
    // (hardware reciprocals, 1 ulp: a relative 2^-23 in the offsets below is
    // far inside the bound's margin)
    const float i0 = __builtin_amdgcn_rcpf((float)g.nc[0]), i1 = __builtin_amdgcn_rcpf((float)g.nc[1]),
                i2 = __builtin_amdgcn_rcpf((float)g.nc[2]);
    // centre -> cell centre, and the half-cell steps along the three cell
    // vectors: centre -> corner (a, b, c) = e + (2a - 1) h0 + (2b - 1) h1 + (2c - 1) h2
    const float d0 = (float)fr[0] - (c0 + 0.5f) * i0, d1 = (float)fr[1] - (c1 + 0.5f) * i1,
                d2 = (float)fr[2] - (c2 + 0.5f) * i2;
    const float ex = H[0] * d0 + H[3] * d1 + H[6] * d2;
    const float ey = H[1] * d0 + H[4] * d1 + H[7] * d2;
    const float ez = H[2] * d0 + H[5] * d1 + H[8] * d2;
    const float h0x = 0.5f * i0 * H[0], h0y = 0.5f * i0 * H[1], h0z = 0.5f * i0 * H[2];
    const float h1x = 0.5f * i1 * H[3], h1y = 0.5f * i1 * H[4], h1z = 0.5f * i1 * H[5];
    const float h2x = 0.5f * i2 * H[6], h2y = 0.5f * i2 * H[7], h2z = 0.5f * i2 * H[8];
    float best = INFINITY;
    if (L.e[0] != 0u) best = __uint_as_float(0xffffffffu - L.e[0]) + sqrtf(ex * ex + ey * ey + ez * ez);
#pragma unroll
    for (int t = 0; t < 8; t++) {
        if (L.e[1 + t] == 0u) continue;
        const float sa = (t & 1) ? 1.0f : -1.0f, sb = (t & 2) ? 1.0f : -1.0f, sc = (t & 4) ? 1.0f : -1.0f;
        const float qx = ex - sa * h0x - sb * h1x - sc * h2x;
        const float qy = ey - sa * h0y - sb * h1y - sc * h2y;
        const float qz = ez - sa * h0z - sb * h1z - sc * h2z;
        best = fminf(best, __uint_as_float(0xffffffffu - L.e[1 + t]) + sqrtf(qx * qx + qy * qy + qz * qz));
    }
    if (!(best < INFINITY)) return INFINITY;
    return (best + __uint_as_float(g.rv_bits)) * (1.0f + 1e-5f) + 1e-2f;
}

// real pass: centre of every solvent molecule (made whole around its first
// atom) wrapped into the cell; near molecules get the Morton key of the cell
template <class T>
__device__ __forceinline__ void b_prep_solvent_tiles(KP P, Bufs B, const T* xv, int bx_, int nbx_, int by_, int nby_)
{
    int64_t total = (int64_t)P.F * P.Mv;
    for (int64_t it = bx_ * (int64_t)blockDim.x + threadIdx.x; it < total; it += (int64_t)nbx_ * blockDim.x) {
        int f = (int)(it / P.Mv);
        int v = (int)(it - (int64_t)f * P.Mv);
        const Geo g = ldc_geo(B.geo + f);
        double c[3];
        mol_centre(g, xv + (int64_t)f * P.Nv * 3, (int64_t)v * P.nv, P.nv, c);
        double fr[3], n[3];
        wrap_point(g, c[0], c[1], c[2], fr, n);
        int cell, c0, c1, c2;
        cell_of_frac(g, fr[0], fr[1], fr[2], cell, c0, c1, c2);
        int64_t kb = (int64_t)f * P.NP * P.Mv + v;
        if (!B.near[(int64_t)f * P.cap + cell]) { B.mkey[kb] = -1; continue; }
        const int key = morton_key3(g, c0, c1, c2);
        B.mkey[kb] = key;
        B.mrank[kb] = atomicAdd(&B.mcnt[(int64_t)f * P.NP * (P.kcap + 1) + key], 1);
    }
}

template <class T>
__global__ void k_prep_solvent_tiles(KP P, Bufs B, const T* xv)
{
    b_prep_solvent_tiles<T>(P, B, xv, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// real pass: atoms of the near molecules in sorted order, FP32, in the frame of
// the centre's cell
template <class T>
__device__ __forceinline__ void b_fill_real(KP P, Bufs B, const T* xv, int bx_, int nbx_, int by_, int nby_)
{
    int64_t total = (int64_t)P.F * P.Mv;
    for (int64_t it = bx_ * (int64_t)blockDim.x + threadIdx.x; it < total; it += (int64_t)nbx_ * blockDim.x) {
        int f = (int)(it / P.Mv);
        int pos = (int)(it - (int64_t)f * P.Mv);
        const int64_t fp = (int64_t)f * P.NP;
        if (pos >= B.nnear[fp]) continue;
        const int v = B.msorted[fp * P.Mv + pos];
        const Geo g = ldc_geo(B.geo + f);
        const T* x = xv + (int64_t)f * P.Nv * 3;
        int64_t a0 = (int64_t)v * P.nv;
        double c[3];
        mol_centre(g, x, a0, P.nv, c);
        double fr[3], n[3];
        wrap_point(g, c[0], c[1], c[2], fr, n);
        const double* H = g.cell.H;
        double tx = H[0] * n[0] + H[3] * n[1] + H[6] * n[2];
        double ty = H[1] * n[0] + H[4] * n[1] + H[7] * n[2];
        double tz = H[2] * n[0] + H[5] * n[1] + H[8] * n[2];
        double x0, y0, z0;
        load3(x, a0, x0, y0, z0);
        // the first atom's w carries the molecule's nearest-solute bound
        int cell, c0, c1, c2;
        cell_of_frac(g, fr[0], fr[1], fr[2], cell, c0, c1, c2);
        const float ub = molecule_bound(g, bound_loads(P, B, g, f, cell, c0, c1, c2), c0, c1, c2, fr);
        float4* dst = B.vw + (int64_t)f * P.Nv + (int64_t)pos * P.nv;
        for (int a = 0; a < P.nv; a++) {
            double xa, ya, za, ox, oy, oz;
            load3(x, a0 + a, xa, ya, za);
            cm::minimage(g.cell, xa - x0, ya - y0, za - z0, &ox, &oy, &oz);
            dst[a] = mk4((float)(x0 + ox - tx), (float)(y0 + oy - ty), (float)(z0 + oz - tz), a == 0 ? ub : 0.f);
        }
    }
}

template <class T>
__global__ void k_fill_real(KP P, Bufs B, const T* xv)
{
    b_fill_real<T>(P, B, xv, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// exclusive scan of the cell counts of passes [p0, p0+np) of every frame (one
// block per (frame, pass)); writes the number of near molecules of the pass
// (the counts stay: the scatter counts them back down to zero)
__device__ __forceinline__ void b_scan_rows(KP P, Bufs B, int p0, int np, int bx_, int nbx_, int by_, int nby_)
{
    __shared__ int wsum[16];
    __shared__ int carry;
    const int f = bx_ / np, pp = p0 + (int)(bx_ - (bx_ / np) * np);
    const int64_t row = (int64_t)f * P.NP + pp;
    int* c = B.mcnt + row * (P.kcap + 1);
    int* st = B.mstart + row * (P.kcap + 1);
    const int n = P.kcap;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // four consecutive keys per thread, the threads' sums scanned over the
    // wavefront (DPP) and across the block's wavefronts
    for (int base = 0; base < n; base += 4 * blockDim.x) {
        const int i0 = base + 4 * threadIdx.x;
        int v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = i0 + u < n ? c[i0 + u] : 0;
        const int tsum = v[0] + v[1] + v[2] + v[3];
        const int incl = wave_incl_scan(tsum);
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w = 0; w < nw; w++) { if (w < wid) woff += wsum[w]; tot += wsum[w]; }
        int e = carry + woff + incl - tsum;
        // starts of empty keys are never read (the scatter reads the keys it
        // holds molecules of): most keys are empty, so only the others are written
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (i0 + u < n && v[u] != 0) st[i0 + u] = e;
            e += v[u];
        }
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        st[n] = carry;
        B.nnear[row] = carry;
    }
}

__global__ void __launch_bounds__(1024) k_scan_rows(KP P, Bufs B, int p0, int np)
{
    b_scan_rows(P, B, p0, np, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// molecules of passes [p0, p0+np) sorted by cell
__device__ __forceinline__ void b_mol_scatter(KP P, Bufs B, int p0, int np, int bx_, int nbx_, int by_, int nby_)
{
    // grid: x over the molecules, y over the (frame, pass) pairs
    const int v = bx_ * blockDim.x + threadIdx.x;
    if (v >= P.Mv) return;
    for (int fq = by_; fq < P.F * np; fq += nby_) {
        const int f = fq / np;
        const int p = p0 + fq - f * np;
        int64_t fp = (int64_t)f * P.NP + p;
        int key = B.mkey[fp * P.Mv + v];
        if (key < 0) continue;
        // the key's range starts where the scan put it; the molecule's rank
        // within the key was taken when the key was counted
        const int pos = B.mstart[fp * (P.kcap + 1) + key] + B.mrank[fp * P.Mv + v];
        B.msorted[fp * P.Mv + pos] = v;
    }
}

__global__ void k_mol_scatter(KP P, Bufs B, int p0, int np)
{
    b_mol_scatter(P, B, p0, np, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// chunk offsets of the passes [p0, p0+np) of every frame (one block)
__device__ __forceinline__ void b_chunk_table(KP P, Bufs B, int p0, int np, int bx_, int nbx_, int by_, int nby_)
{
    __shared__ int wsum[16];
    __shared__ int carry;
    int n = P.F * np;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int base = 0; base < n; base += blockDim.x) {
        int idx = base + threadIdx.x;
        int v = 0;
        if (idx < n) {
            int f = idx / np, p = p0 + idx - f * np;
            int nn = B.nnear[(int64_t)f * P.NP + p];
            v = (nn + P.mpt - 1) / P.mpt;
        }
        int incl = v;
        for (int o = 1; o < 64; o <<= 1) {
            int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w = 0; w < nw; w++) { if (w < wid) woff += wsum[w]; tot += wsum[w]; }
        if (idx < n) B.chunk_base[idx] = carry + woff + incl - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        B.chunk_base[n] = carry;
        B.chunk_base[n + 1] = 0;   // the evaluation's chunk counter
    }
}

__global__ void __launch_bounds__(1024) k_chunk_table(KP P, Bufs B, int p0, int np)
{
    b_chunk_table(P, B, p0, np, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// randomize_solvent! (src/mddf.jl:65-88), first half: the new centre of every
// slot; slots whose centre falls in a near cell get the Morton key of the cell
// (the others cannot have an atom within the cutoff of the solute)
__device__ __forceinline__ void b_place_keys(KP P, Bufs B, int bx_, int nbx_, int by_, int nby_)
{
    // grid: x over the slots, y over the (frame, pass) pairs
    const int v = bx_ * blockDim.x + threadIdx.x;
    if (v >= P.Mv) return;
    for (int fk = by_; fk < P.F * P.K; fk += nby_) {
        const int f = fk / P.K;
        const int k = fk - f * P.K;
        const Geo g = ldc_geo(B.geo + f);
        int64_t fp = (int64_t)f * P.NP + 1 + k;
        cm::u32x4 r = cm::philox4x32_10({(uint32_t)v, (uint32_t)k, g.frame, cm::TAG_SLOT}, P.key0, P.key1);
        int cell, c0, c1, c2;
        cell_of_draws(g, r.y, r.z, r.w, cell, c0, c1, c2);
        if (!B.near[(int64_t)f * P.cap + cell]) { B.mkey[fp * P.Mv + v] = -1; continue; }
        const int key = morton_key3(g, c0, c1, c2);
        B.mkey[fp * P.Mv + v] = key;
        B.mrank[fp * P.Mv + v] = atomicAdd(&B.mcnt[fp * (P.kcap + 1) + key], 1);
    }
}

__global__ void k_place_keys(KP P, Bufs B)
{
    b_place_keys(P, B, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// second half: the placement of the near slots, written in sorted order (FP64
// for the exact resolve, FP32 for the evaluation)
template <class TV>
__device__ __forceinline__ void b_place_near(KP P, Bufs B, const TV* xv, int bx_, int nbx_, int by_, int nby_)
{
    // grid: x over the sorted positions, y over the (frame, pass) pairs
    const int pos = bx_ * blockDim.x + threadIdx.x;
    if (pos >= P.Mv) return;
    for (int fk = by_; fk < P.F * P.K; fk += nby_) {
        const int f = fk / P.K;
        const int k = fk - f * P.K;
        int64_t fp = (int64_t)f * P.NP + 1 + k;
        if (pos >= ldg(B.nnear + fp)) continue;
        const int v = ldg(B.msorted + fp * P.Mv + pos);
        const Geo g = ldc_geo(B.geo + f);
        cm::u32x4 r = cm::philox4x32_10({(uint32_t)v, (uint32_t)k, g.frame, cm::TAG_SLOT}, P.key0, P.key1);
        cm::u32x4 r2 = cm::philox4x32_10({(uint32_t)v, (uint32_t)k, g.frame, cm::TAG_SLOT + 1u}, P.key0, P.key1);
        int64_t lk = (int64_t)f * P.kb;
        int nb = ldg(B.nbulk + lk);
        int64_t jmol = nb > 0 ? (int64_t)ldg(B.blist + lk * P.Mv + cm::draw_index(r.x, nb)) : cm::draw_index(r.x, P.Mv);
        MolSrc<TV> src(xv + (int64_t)f * P.Nv * 3 + jmol * P.nv * 3);
        MolDst dst(B.P + (fk * P.Nv + (int64_t)pos * P.nv) * 3, B.vw_r + fk * P.Nv + (int64_t)pos * P.nv);
        double uc[3] = {cm::u01(r.y), cm::u01(r.z), cm::u01(r.w)};
        double ua[3] = {cm::u01(r2.x), cm::u01(r2.y), cm::u01(r2.z)};
        // the placed centre is H uc (a rigid copy of a molecule of the frame):
        // the first atom's w carries the molecule's nearest-solute bound
        int cell, c0, c1, c2;
        cell_of_frac(g, uc[0], uc[1], uc[2], cell, c0, c1, c2);
        const float ub = molecule_bound(g, bound_loads(P, B, g, f, cell, c0, c1, c2), c0, c1, c2, uc);
        cm::random_move(src, dst, P.nv, P.iref, g.cell, uc, ua);
        dst.set_w0(ub);
    }
}

template <class TV>
__global__ void k_place_near(KP P, Bufs B, const TV* xv)
{
    b_place_near<TV>(P, B, xv, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

// lexicographic minimum of two exact records, OR of the reference flags
__device__ __forceinline__ void rec_merge(Rec& r, const Rec& o)
{
    if (o.within && (!r.within || o.d < r.d || (o.d == r.d && (o.i < r.i || (o.i == r.i && o.j < r.j))))) {
        r.within = 1; r.i = o.i; r.j = o.j; r.d = o.d;
    }
    if (o.ref) { r.ref = 1; r.dref = fmin(r.dref, o.dref); }
}

__device__ __forceinline__ Rec rec_shfl_xor(const Rec& r, int o)
{
    Rec q;
    q.within = __shfl_xor(r.within, o, 64);
    q.i = __shfl_xor(r.i, o, 64);
    q.j = (int64_t)__shfl_xor((long long)r.j, o, 64);
    q.d = __shfl_xor(r.d, o, 64);
    q.ref = __shfl_xor(r.ref, o, 64);
    q.dref = __shfl_xor(r.dref, o, 64);
    return q;
}

// Column reduction of a running minimum over the C lanes of a row: DPP
// quad permutations for the partners 1 and 2 lanes away (no LDS traffic),
// a lane shuffle beyond.
template <int O>
__device__ __forceinline__ int xor_lane(int x)
{
    if constexpr (O == 1) return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xf, 0xf, false);   // quad_perm 1,0,3,2
    else if constexpr (O == 2) return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xf, 0xf, false);   // quad_perm 2,3,0,1
    else return __shfl_xor(x, O, 64);
}

template <int O>
__device__ __forceinline__ void col_merge(Best& b)
{
    const float b2 = __int_as_float(xor_lane<O>(__float_as_int(b.best)));
    const float s2 = __int_as_float(xor_lane<O>(__float_as_int(b.second)));
    const int i2 = xor_lane<O>(b.bi);
    const float ns = fminf(fmaxf(b.best, b2), fminf(b.second, s2));
    if (b2 < b.best) b.bi = i2;
    b.best = fminf(b.best, b2);
    b.second = ns;
}

template <int C>
__device__ __forceinline__ void col_reduce(Best& b)
{
    if constexpr (C >= 2) col_merge<1>(b);
    if constexpr (C >= 4) col_merge<2>(b);
    if constexpr (C >= 8) col_merge<4>(b);
}

// Minimum distances of the solvent molecules of one chunk per wavefront
// (64 threads).  Lane = (row, column); the C columns of a row split the
// staged solute atoms between them.  NVM = 0: a row is one solvent atom and
// the atoms of a molecule are combined across rows.  NVM > 0 (molecules of
// NVM atoms, e.g. water): a row is one molecule, each lane evaluates every
// staged atom against all NVM atoms it holds in registers (one LDS read per
// NVM pairs) and combines them in the lane.
template <class TU, class TV, bool RANDOM, int C, int NVM, bool VERIFY>
__device__ __forceinline__ void eval_tiles_chunks(const KP& P, const Bufs& B, const TU* xu, const TV* xv, int ch0, int ch1,
                                                  int chs, float4* stg, float* stw, const int* cbs, bool cb_lds,
                                                  int* ctr)
{
    // chunks of one job below ch1: ch0, ch0 + chs, ... below tdyn, then
    // (with a counter ctr, zero at the start of the launch) the chunks from
    // tdyn on in the order the workgroups ask for them, so the ones that
    // finish early take more of the last chunks.  The chunks' costs vary
    // with their staged atoms, and a fixed stride alone leaves a tail of
    // workgroups still busy; requests go to one device-wide counter, so
    // only the last DYN_PCT per cent are handed out that way, a few
    // consecutive chunks per request.
    // LDS: the staging buffer stg (P.tcap float4: atom, index; relative to
    // the chunk centre when NVM > 0), with NVM > 0 the squared norms stw of
    // the staged atoms, and (cb_lds) the job's chunk table cbs
    const int tcap = P.tcap;
    // staging ranks per lane per step (fewer with a molecule per lane, whose
    // evaluation holds more registers)
    constexpr int SUk = NVM > 0 ? 2 : SU;
    const int lane = threadIdx.x;
    const int row = lane / C, col = lane - (lane / C) * C;
    const int p0 = RANDOM ? 1 : 0;
    const int npass = RANDOM ? P.K : 1;
    const int nfp = P.F * npass;
    ull* gu = RANDOM ? B.cnt_gu_r : B.cnt_gu;
    ull* gv = RANDOM ? B.cnt_gv_r : B.cnt_gv;
    ull* gmd = RANDOM ? B.cnt_md_r : B.cnt_md;
    ull* grdf = RANDOM ? B.cnt_rdf_r : B.cnt_rdf;
    const int ndyn = (int)((int64_t)ch1 * DYN_PCT / 100);
    const int tdyn = ctr && ch1 >= DYN_MIN * chs ? ch1 - ndyn : ch1;
    const int dgrab = max(1, ndyn / (chs * DYN_GRABS));   // chunks per request
    int dnext = 0, dend = 0;
    auto next_chunk = [&](int c) {
        if (c < tdyn || tdyn == ch1) return c;
        if (dnext < dend) return dnext++;
        int g = 0;
        if (lane == 0) g = atomicAdd(ctr, dgrab);
        dnext = tdyn + __builtin_amdgcn_readlane(g, 0);
        dend = dnext + dgrab;
        return dnext++;
    };
    for (int ch = next_chunk(ch0); ch < ch1; ch = next_chunk(ch + chs)) {
        int lo_ = 0, hi_ = nfp;
        while (hi_ - lo_ > 1) {
            int mid = (lo_ + hi_) >> 1;
            if ((cb_lds ? cbs[mid] : B.chunk_base[mid]) <= ch) lo_ = mid; else hi_ = mid;
        }
        const int fq = lo_;
        const int f = fq / npass, pidx = fq - (fq / npass) * npass;
        const int p = p0 + pidx;
        const int64_t fp = (int64_t)f * P.NP + p;
        const Geo& g = B.geo[f];
        const int m0 = (ch - (cb_lds ? cbs[fq] : B.chunk_base[fq])) * P.mpt;
        const int ml = NVM ? row : row / P.nv, a = NVM ? 0 : row - (row / P.nv) * P.nv;
        // the molecule's loads are issued together with the near count that
        // decides whether the lane has a molecule (positions past the count
        // read a valid slot and are discarded)
        const int mpos = min(m0 + ml, P.Mv - 1);
        const int nm = min(P.mpt, B.nnear[fp] - m0);
        const int vld = B.msorted[fp * P.Mv + mpos];
        const float4* vwb = RANDOM ? B.vw_r + ((int64_t)f * P.K + pidx) * P.Nv : B.vw + (int64_t)f * P.Nv;
        const int64_t pbase = (int64_t)mpos * P.nv;   // sorted position of the molecule
        float4 y = mk4(0.f, 0.f, 0.f, 0.f);
        float4 ym[NVM > 0 ? NVM : 1];   // molecule per lane: its atoms
        if constexpr (NVM > 0) {
#pragma unroll
            for (int q = 0; q < NVM; q++) ym[q] = vwb[pbase + q];
        } else {
            y = vwb[pbase + a];
        }
        const bool active = ml < nm;
        const int v = active ? vld : 0;
        float alx, aly, alz, ahx, ahy, ahz;   // box of the lane's atoms
        if constexpr (NVM > 0) {
#pragma unroll
            for (int q = 0; q < NVM; q++) if (!active) ym[q] = mk4(0.f, 0.f, 0.f, 0.f);
            alx = ahx = ym[0].x; aly = ahy = ym[0].y; alz = ahz = ym[0].z;
#pragma unroll
            for (int q = 1; q < NVM; q++) {
                alx = fminf(alx, ym[q].x); aly = fminf(aly, ym[q].y); alz = fminf(alz, ym[q].z);
                ahx = fmaxf(ahx, ym[q].x); ahy = fmaxf(ahy, ym[q].y); ahz = fmaxf(ahz, ym[q].z);
            }
        } else {
            if (!active) y = mk4(0.f, 0.f, 0.f, 0.f);
            alx = ahx = y.x; aly = ahy = y.y; alz = ahz = y.z;
        }
        const float lx = wave_min(active ? alx : INFINITY), ly = wave_min(active ? aly : INFINITY),
                    lz = wave_min(active ? alz : INFINITY);
        const float hx = wave_max(active ? ahx : -INFINITY), hy = wave_max(active ? ahy : -INFINITY),
                    hz = wave_max(active ? ahz : -INFINITY);
        const float rs_full = g.rstage;
        // NVM > 0: distances relative to the chunk centre c.  With y' = y - c
        // and s' = s - c, |y' - s'|^2 = (|s'|^2 - 2 y'.s') + |y'|^2: three FMAs
        // per pair against the staged (s', |s'|^2), |y'|^2 added once per atom.
        // With u = 2^-24, Y >= |y'| and S >= |s'| for every staged atom, the
        // rounding of the c subtractions moves y' - s' by at most u (Y + S),
        // so d^2 by at most 2 u (Y + S) d (d <= cutsys + 0.5 where it
        // matters); the norms (3 u S^2, 3 u Y^2), the FMA chain (3 u (S^2 +
        // 2 Y S)) and the last sum (u (Y + S)^2) add at most 7 u (Y + S)^2.
        // Both widen the frame's guard for this chunk.
        // The molecule's atoms are kept as -2 y' (exact scaling of y').
        // The per-chunk values are wave-uniform and held in scalar registers.
        float ccx = 0.f, ccy = 0.f, ccz = 0.f;
        float Gcf = (float)g.guard, cut2_hi_c = g.cut2_hi, amb_c = g.amb, G32c = g.guard32;
        if constexpr (NVM > 0) {
            ccx = sgpr_f(0.5f * (lx + hx)); ccy = sgpr_f(0.5f * (ly + hy)); ccz = sgpr_f(0.5f * (lz + hz));
            const double ex = (double)hx - lx, ey = (double)hy - ly, ez = (double)hz - lz;
            // (+ the rounding of c itself, at most 2^-24 |c| per component)
            const double Yb = 0.5 * sqrt(ex * ex + ey * ey + ez * ez) * (1.0 + 1e-6) +
                              1e-6 * ((double)fabsf(ccx) + fabsf(ccy) + fabsf(ccz)) + 1e-5;
            const double Sb = Yb + (double)rs_full * (1.0 + 1e-5) + 1e-4;
            const double Gc = g.guard + 0x1p-24 * (Yb + Sb) * (2.0 * (P.cutsys + 0.5) + 7.0 * (Yb + Sb)) + 1e-9;
            Gcf = sgpr_f(nextafterf((float)Gc, INFINITY));
            cut2_hi_c = sgpr_f(nextafterf((float)(P.cut2 + Gc), INFINITY));
            amb_c = sgpr_f(nextafterf((float)(2.0 * Gc) + 8.0f * (float)(halfulp32((double)cut2_hi_c) * 2.0), INFINITY));
            G32c = sgpr_f(nextafterf((float)(Gc + (P.cut2 + Gc + 1.0) * 0x1p-21), INFINITY));
#pragma unroll
            for (int q = 0; q < NVM; q++)
                ym[q] = mk4(-2.0f * (ym[q].x - ccx), -2.0f * (ym[q].y - ccy), -2.0f * (ym[q].z - ccz), ym[q].w);
        }
        // Staging radius.  Each molecule carries an upper bound U on its
        // atoms' nearest solute distance (molecule_bound).  A solute atom that
        // can decide anything for an atom p of the chunk (its FP64 nearest,
        // or within the FP32 decision band of it) is within
        // sqrt(U^2 + 2 G + amb) of p, hence of the AABB; staging out to
        // sqrt(U^2 + 2 amb + 4 G) plus a slack over FP32 rounding keeps every
        // such atom, and the farther ones cannot change a decision.
        float rs = rs_full;
        {
            const float um = wave_max(active ? (NVM > 0 ? ym[0].w : y.w) : 0.f);
            if (um < rs_full)
                rs = sgpr_f(fminf(rs_full, sqrtf(fmaf(um, um, 2.0f * amb_c + 4.0f * Gcf)) * (1.0f + 1e-5f) + 1e-2f));
        }
        const float rs2 = rs * rs;
        int ilo[3], ihi[3];
        cell_range(g, lx, ly, lz, hx, hy, hz, rs, ilo, ihi);
        const int nx = ihi[0] - ilo[0] + 1, ny = ihi[1] - ilo[1] + 1, nz = ihi[2] - ilo[2] + 1;
        const int ncl = nx * ny * nz;
        const int* st = B.start + (int64_t)f * (P.cap + 1);
        const float4* srt = B.sorted + (int64_t)f * B.nitems;
        Best bb;
        bb.init();
        Best bq[NVM > 0 ? NVM : 1];   // molecule per lane: each atom's running minimum
#pragma unroll
        for (int q = 0; q < (NVM > 0 ? NVM : 1); q++) bq[q].init();
        int rounds = 0;
        int nst = 0;   // staged atoms (uniform)
        const unsigned long long lt_mask = (1ull << lane) - 1ull;
        auto evaluate = [&](int n) {
            __syncthreads();
            if constexpr (NVM > 0) {
                // |s'|^2 - 2 y'.s' per atom of the molecule (|y'|^2 added after
                // the last round); four staged atoms per step, loads first
                int k = col;
                for (; k + 3 * C < n; k += 4 * C) {
                    float4 sv[4];
                    float wv[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) { sv[u] = stg[k + u * C]; wv[u] = stw[k + u * C]; }
#pragma unroll
                    for (int u = 0; u < 4; u++)
#pragma unroll
                        for (int q = 0; q < NVM; q++)
                            bq[q].offer(fmaf(ym[q].x, sv[u].x, fmaf(ym[q].y, sv[u].y, fmaf(ym[q].z, sv[u].z, wv[u]))),
                                        __float_as_int(sv[u].w), 0);
                }
                for (; k < n; k += C) {
                    const float4 s0 = stg[k];
                    const float w0 = stw[k];
#pragma unroll
                    for (int q = 0; q < NVM; q++)
                        bq[q].offer(fmaf(ym[q].x, s0.x, fmaf(ym[q].y, s0.y, fmaf(ym[q].z, s0.z, w0))), __float_as_int(s0.w), 0);
                }
                __syncthreads();
                return;
            }
            int k = col;
            for (; k + 3 * C < n; k += 4 * C) {
                float4 s0 = stg[k], s1 = stg[k + C], s2 = stg[k + 2 * C], s3 = stg[k + 3 * C];
                float dx, dy, dz;
                dx = y.x - s0.x; dy = y.y - s0.y; dz = y.z - s0.z;
                bb.offer(fmaf(dx, dx, fmaf(dy, dy, dz * dz)), __float_as_int(s0.w), 0);
                dx = y.x - s1.x; dy = y.y - s1.y; dz = y.z - s1.z;
                bb.offer(fmaf(dx, dx, fmaf(dy, dy, dz * dz)), __float_as_int(s1.w), 0);
                dx = y.x - s2.x; dy = y.y - s2.y; dz = y.z - s2.z;
                bb.offer(fmaf(dx, dx, fmaf(dy, dy, dz * dz)), __float_as_int(s2.w), 0);
                dx = y.x - s3.x; dy = y.y - s3.y; dz = y.z - s3.z;
                bb.offer(fmaf(dx, dx, fmaf(dy, dy, dz * dz)), __float_as_int(s3.w), 0);
            }
            for (; k < n; k += C) {
                float4 s0 = stg[k];
                float dx = y.x - s0.x, dy = y.y - s0.y, dz = y.z - s0.z;
                bb.offer(fmaf(dx, dx, fmaf(dy, dy, dz * dz)), __float_as_int(s0.w), 0);
            }
            __syncthreads();
        };
        // Stage the solute atoms of the cells in range, 64 cells at a time: the
        // atoms of those cells are ranked (wave scan of the counts), each lane
        // loads the atoms of its ranks, and the ones within rs of the AABB are
        // appended at ballot/popcount offsets.
        // Staging walks rows of cells (x fastest in memory): a slot is the part
        // of one (y, z) row of the index box that lies in one periodic image
        // along x, so its atoms are one contiguous range of the sorted list.
        // With a diagonal cell, rows farther than rs from the AABB are
        // dropped and each row is trimmed to the x range within rs (plus a
        // slack over FP32 rounding); the per-atom AABB test below is exact.
        int q0;
        wrap_index(ilo[0], g.nc[0], q0);
        int q1;
        wrap_index(ihi[0], g.nc[0], q1);
        const int npc = q1 - q0 + 1;
        const int nslots = ny * nz * npc;
        const float rsl = rs + 1e-3f, rsl2 = rsl * rsl;
        const float hcx = g.H32[0] / g.nc[0], hcy = g.H32[4] / g.nc[1], hcz = g.H32[8] / g.nc[2];
        for (int sb = 0; sb < nslots; sb += 64) {
            const int si = sb + lane;
            int cstart = 0, cnt = 0;
            float tx = 0.f, ty = 0.f, tz = 0.f;
            if (si < nslots) {
                const int r = si / npc, pc = si - r * npc;
                const int iy = ilo[1] + r % ny, iz = ilo[2] + r / ny;
                int x0 = ilo[0], x1 = ihi[0];
                bool keep = true;
                if (g.ortho) {
                    const float dy = fmaxf(fmaxf(iy * hcy - hy, ly - (iy + 1) * hcy), 0.f);
                    const float dz = fmaxf(fmaxf(iz * hcz - hz, lz - (iz + 1) * hcz), 0.f);
                    const float r2 = rsl2 - dy * dy - dz * dz;
                    if (r2 < 0.f) {
                        keep = false;
                    } else {
                        const float rx = sqrtf(r2) + 1e-3f;
                        x0 = max(x0, (int)floorf((lx - rx) / hcx));
                        x1 = min(x1, (int)floorf((hx + rx) / hcx));
                    }
                }
                const int mx = q0 + pc;
                x0 = max(x0, mx * g.nc[0]);
                x1 = min(x1, mx * g.nc[0] + g.nc[0] - 1);
                if (keep && x0 <= x1) {
                    int my, mz;
                    const int cy = wrap_index(iy, g.nc[1], my), cz = wrap_index(iz, g.nc[2], mz);
                    const int crow = (cz * g.nc[1] + cy) * g.nc[0] - mx * g.nc[0];
                    cstart = st[crow + x0];
                    cnt = st[crow + x1 + 1] - cstart;
                    shift_of(g, mx, my, mz, tx, ty, tz);
                }
            }
            // The atoms of the rows are staged in segments of up to SEG
            // consecutive atoms: each lane takes one segment per step (owner
            // row found by a search over the wave's segment prefix, once per
            // SEG atoms), loads its atoms SUk at a time with the loads in
            // flight together, and appends the ones within rs of the AABB at
            // ballot/popcount offsets.  The staged set does not depend on the
            // order (FP32 ties between candidates go to the FP64 resolve).
            const int nseg = (cnt + SEG - 1) / SEG;
            const int sincl = wave_incl_scan(nseg);
            const int sexcl = sincl - nseg;
            const int NS = __builtin_amdgcn_readlane(sincl, 63);
            for (int s0 = 0; s0 < NS; s0 += 64) {
                const int sg = s0 + lane;
                int j = 0;
#pragma unroll
                for (int step = 32; step >= 1; step >>= 1) {
                    const int e = __shfl(sexcl, (j + step) & 63, 64);
                    j = (j + step < 64 && e <= sg) ? j + step : j;
                }
                const int sj = __shfl(cstart, j, 64), cj = __shfl(cnt, j, 64), ej = __shfl(sexcl, j, 64);
                const float ox = __shfl(tx, j, 64), oy = __shfl(ty, j, 64), oz = __shfl(tz, j, 64);
                const int k0 = (sg - ej) * SEG;
                const int n = sg < NS ? min(SEG, cj - k0) : 0;
                const int abase = sj + k0;
                const int nmax = wave_max_i(n);
                for (int kk = 0; kk < nmax; kk += SUk) {
                    float4 q[SUk];
#pragma unroll
                    for (int u = 0; u < SUk; u++) q[u] = srt[kk + u < n ? abase + kk + u : 0];
#pragma unroll
                    for (int u = 0; u < SUk; u++) q[u] = mk4(q[u].x + ox, q[u].y + oy, q[u].z + oz, q[u].w);
#pragma unroll
                    for (int u = 0; u < SUk; u++) {
                        if (kk + u >= nmax) break;
                        if (nst + 64 > tcap) {
                            evaluate(nst);
                            rounds++;
                            nst = 0;
                        }
                        const float dx = fmaxf(fmaxf(lx - q[u].x, q[u].x - hx), 0.f);
                        const float dy = fmaxf(fmaxf(ly - q[u].y, q[u].y - hy), 0.f);
                        const float dz = fmaxf(fmaxf(lz - q[u].z, q[u].z - hz), 0.f);
                        const bool pass = kk + u < n && dx * dx + dy * dy + dz * dz <= rs2;
                        const unsigned long long m = __ballot(pass);
                        if (pass) {
                            const int slot = nst + __popcll(m & lt_mask);
                            if constexpr (NVM > 0) {
                                const float sx_ = q[u].x - ccx, sy_ = q[u].y - ccy, sz_ = q[u].z - ccz;
                                stg[slot] = mk4(sx_, sy_, sz_, q[u].w);
                                stw[slot] = fmaf(sx_, sx_, fmaf(sy_, sy_, sz_ * sz_));
                            } else {
                                stg[slot] = q[u];
                            }
                        }
                        nst += __popcll(m);
                    }
                }
            }
        }
        const bool one_round = rounds == 0;
        const int last_n = nst;   // the last staged batch stays in LDS until the next chunk
        if (last_n > 0) evaluate(last_n);
        float mb, ms, rb, rsec;
        int mi, ma, ri;
        if constexpr (NVM > 0) {
            // + |y'|^2; columns -> each atom's minimum; atoms -> the molecule, in the lane
#pragma unroll
            for (int q = 0; q < NVM; q++) {
                const float yy2 = 0.25f * fmaf(ym[q].x, ym[q].x, fmaf(ym[q].y, ym[q].y, ym[q].z * ym[q].z));
                bq[q].best += yy2;
                bq[q].second += yy2;
            }
#pragma unroll
            for (int q = 0; q < NVM; q++) col_reduce<C>(bq[q]);
            mb = bq[0].best; ms = bq[0].second; mi = bq[0].bi; ma = 0;
            rb = bq[0].best; rsec = bq[0].second; ri = bq[0].bi;
#pragma unroll
            for (int q = 1; q < NVM; q++) {
                const float ns = fminf(fmaxf(mb, bq[q].best), fminf(ms, bq[q].second));
                if (bq[q].best < mb) { mi = bq[q].bi; ma = q; }
                mb = fminf(mb, bq[q].best);
                ms = ns;
                if (q == P.iref) { rb = bq[q].best; rsec = bq[q].second; ri = bq[q].bi; }
            }
        } else {
            // columns of a row -> the row's atom
            col_reduce<C>(bb);
            // atoms of a molecule -> the molecule (leader lane: atom 0, column 0)
            const int mrow0 = ml * P.nv;
            mb = bb.best; ms = bb.second;
            mi = bb.bi; ma = 0;
            for (int q = 1; q < P.nv; q++) {
                const int src = ((mrow0 + q) * C) & 63;
                float b2 = __shfl(bb.best, src, 64), s2 = __shfl(bb.second, src, 64);
                int i2 = __shfl(bb.bi, src, 64);
                float ns = fminf(fmaxf(mb, b2), fminf(ms, s2));
                if (b2 < mb) { mi = i2; ma = q; }
                mb = fminf(mb, b2);
                ms = ns;
            }
            const int srcr = ((mrow0 + P.iref) * C) & 63;
            rb = __shfl(bb.best, srcr, 64);
            rsec = __shfl(bb.second, srcr, 64);
            ri = __shfl(bb.bi, srcr, 64);
        }
        const bool leader = active && a == 0 && col == 0;
        const TU* xuf = xu + (int64_t)f * P.Nu * 3;
        const int64_t jbase = (int64_t)v * P.nv;
        auto ysrc = [&](int q, double& y0, double& y1, double& y2) {
            if (RANDOM) {
                const double* pp = B.P + (((int64_t)f * P.K + pidx) * P.Nv + pbase + q) * 3;
                y0 = pp[0]; y1 = pp[1]; y2 = pp[2];
            } else load3(xv + (int64_t)f * P.Nv * 3, jbase + q, y0, y1, y2);
        };
        Rec r = rec_zero();
        bool exh = false;
        bool d32 = false, dref32 = false;   // r.d / r.dref hold FP32 squared distances (bins decided without them)
        if (leader) {
            if (mb <= cut2_hi_c) {
                if (ms - mb > amb_c) {
                    int bn, bk;
                    const int dd = decide32(P, G32c, mb, bn, bk);
                    if (dd != 0) {
                        if (dd == 2) { r.within = 1; r.i = mi; r.j = jbase + ma; r.d = mb; r.bin = bn; r.bulk = bk; d32 = true; }
                    } else {
                        double x0, x1, x2, y0, y1, y2;
                        load3(xuf, mi, x0, x1, x2);
                        ysrc(ma, y0, y1, y2);
                        double d2 = cm::canon_d2(g.cell, x0, x1, x2, y0, y1, y2);
                        if (d2 <= P.cut2) { r.within = 1; r.i = mi; r.j = jbase + ma; r.d = sqrt(d2); }
                    }
                } else exh = true;
            }
            if (rb <= cut2_hi_c) {
                if (rsec - rb > amb_c) {
                    int bn, bk;
                    const int dd = decide32(P, G32c, rb, bn, bk);
                    if (dd != 0) {
                        if (dd == 2) { r.ref = 1; r.dref = rb; r.rbin = bn; dref32 = true; }
                    } else {
                        double x0, x1, x2, y0, y1, y2;
                        load3(xuf, ri, x0, x1, x2);
                        ysrc(P.iref, y0, y1, y2);
                        double d2 = cm::canon_d2(g.cell, x0, x1, x2, y0, y1, y2);
                        if (d2 <= P.cut2) { r.ref = 1; r.dref = sqrt(d2); }
                    }
                } else exh = true;
            }
        }
        // ambiguous molecules: every staged pair within the band, in FP64
        unsigned long long amask = __ballot(exh);
        if (amask && lane == 0) atomicAdd(&B.stats[0], (ull)__popcll(amask));
        while (amask) {
            const int L = __ffsll((long long)amask) - 1;
            amask &= amask - 1ull;
            const float bandm = __shfl(mb, L, 64) + amb_c, bandr = __shfl(rb, L, 64) + amb_c;
            const int mlL = NVM ? L / C : (L / C) / P.nv;
            const int vL = __shfl(v, L, 64);
            const int64_t jb = (int64_t)vL * P.nv;
            const int64_t pb = (int64_t)(m0 + mlL) * P.nv;
            Rec loc = rec_zero();
            for (int q = 0; q < P.nv; q++) {
                const int sl = NVM ? (mlL * C) & 63 : ((mlL * P.nv + q) * C) & 63;
                float4 yq = y;
                if constexpr (NVM > 0) {
#pragma unroll
                    for (int t = 0; t < NVM; t++)
                        if (t == q) yq = ym[t];
                }
                float yx = __shfl(yq.x, sl, 64), yy = __shfl(yq.y, sl, 64), yz = __shfl(yq.z, sl, 64);
                // NVM > 0: back from -2 y' (relative, exact) ...
                if constexpr (NVM > 0) { yx *= -0.5f; yy *= -0.5f; yz *= -0.5f; }
                double Y0, Y1, Y2;
                if (RANDOM) {
                    const double* pp = B.P + (((int64_t)f * P.K + pidx) * P.Nv + pb + q) * 3;
                    Y0 = pp[0]; Y1 = pp[1]; Y2 = pp[2];
                } else load3(xv + (int64_t)f * P.Nv * 3, jb + q, Y0, Y1, Y2);
                const bool isref = q == P.iref;
                if (one_round) {
                    // NVM > 0: y' against the staged atoms relative to the chunk centre
                    for (int k = lane; k < last_n; k += 64) {
                        float4 s = stg[k];
                        float dx = yx - s.x, dy = yy - s.y, dz = yz - s.z;
                        float d2f = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                        if (d2f <= bandm || (isref && d2f <= bandr)) {
                            int i = __float_as_int(s.w);
                            double x0, x1, x2;
                            load3(xuf, i, x0, x1, x2);
                            double d2 = cm::canon_d2(g.cell, x0, x1, x2, Y0, Y1, Y2);
                            if (d2 <= P.cut2) rec_offer(loc, i, jb + q, sqrt(d2), isref);
                        }
                    }
                } else {
                    // staged buffer recycled: scan the cells around the atom out
                    // to the band (an FP32 d2 within the band means an FP64 d2
                    // within band + guard), every pair there in FP64
                    int jlo[3], jhi[3];
                    const float band = isref ? fmaxf(bandm, bandr) : bandm;
                    const float rband = fminf(sqrtf(band + Gcf) + 1e-3f, rs);
                    // ... and to absolute coordinates (rounding within the 1e-3 slack)
                    const float yax = NVM > 0 ? yx + ccx : yx, yay = NVM > 0 ? yy + ccy : yy, yaz = NVM > 0 ? yz + ccz : yz;
                    cell_range(g, yax, yay, yaz, yax, yay, yaz, rband, jlo, jhi);
                    const int mxn = jhi[0] - jlo[0] + 1, myn = jhi[1] - jlo[1] + 1, mzn = jhi[2] - jlo[2] + 1;
                    const int ncq = mxn * myn * mzn;
                    for (int ci = lane; ci < ncq; ci += 64) {
                        int ix = jlo[0] + ci % mxn, iy = jlo[1] + (ci / mxn) % myn, iz = jlo[2] + ci / (mxn * myn);
                        int mx, my, mz;
                        int cx = wrap_index(ix, g.nc[0], mx), cy = wrap_index(iy, g.nc[1], my), cz = wrap_index(iz, g.nc[2], mz);
                        int cell = (cz * g.nc[1] + cy) * g.nc[0] + cx;
                        for (int pq = st[cell]; pq < st[cell + 1]; pq++) {
                            int i = __float_as_int(srt[pq].w);
                            double x0, x1, x2;
                            load3(xuf, i, x0, x1, x2);
                            double d2 = cm::canon_d2(g.cell, x0, x1, x2, Y0, Y1, Y2);
                            if (d2 <= P.cut2) rec_offer(loc, i, jb + q, sqrt(d2), isref);
                        }
                    }
                }
            }
            for (int o = 32; o >= 1; o >>= 1) rec_merge(loc, rec_shfl_xor(loc, o));
            if (lane == L) { r = loc; d32 = dref32 = false; }
        }
        if (leader) {
            rec_complete(P, r);
            hist_record_global(P, gmd, grdf, gu, gv, r);
            if (!RANDOM && P.K > 0) B.bflag[(int64_t)f * P.kb * P.Mv + v] = inbulk(P, r) ? 1 : 0;
            if constexpr (VERIFY) {
                // verification: the record with its FP64 distances (the
                // argmin pairs are exact; the distances of records decided
                // from FP32 bounds are formed here)
                VRec o;
                o.within = r.within; o.i = r.within ? r.i : -1; o.j = r.within ? r.j : -1;
                o.d = r.within ? r.d : INFINITY;
                o.ref = r.ref; o.pad = 0; o.dref = r.ref ? r.dref : INFINITY;
                double x0, x1, x2, y0, y1, y2;
                if (r.within && d32) {
                    load3(xuf, r.i, x0, x1, x2);
                    ysrc((int)(r.j - jbase), y0, y1, y2);
                    o.d = sqrt(cm::canon_d2(g.cell, x0, x1, x2, y0, y1, y2));
                }
                if (r.ref && dref32) {
                    load3(xuf, ri, x0, x1, x2);
                    ysrc(P.iref, y0, y1, y2);
                    o.dref = sqrt(cm::canon_d2(g.cell, x0, x1, x2, y0, y1, y2));
                }
                B.vrec[fp * P.Mv + v] = o;
            }
        }
        __syncthreads();
    }
}

// Minimum distances of the solvent molecules of one chunk per wavefront
// (64 threads), over every chunk of one job's launch (persistent grid).
template <class TU, class TV, bool RANDOM, int C, int NVM, bool VERIFY>
__global__ void __launch_bounds__(64, 3) k_eval_tiles(KP P, Bufs B, const TU* xu, const TV* xv)
{
    // dynamic LDS: staging buffer, squared norms (NVM > 0), then the chunk
    // table of this launch when it has at most CBCAP entries
    extern __shared__ float4 dyn[];
    float4* stg = dyn;
    float* stw = (float*)(dyn + P.tcap);
    int* cbs = (int*)(stw + (NVM > 0 ? P.tcap : 0));
    const int nfp = P.F * (RANDOM ? P.K : 1);
    const bool cb_lds = nfp + 1 <= CBCAP;
    if (cb_lds)
        for (int k = threadIdx.x; k <= nfp; k += 64) cbs[k] = B.chunk_base[k];
    __syncthreads();
    eval_tiles_chunks<TU, TV, RANDOM, C, NVM, VERIFY>(P, B, xu, xv, blockIdx.x, B.chunk_base[nfp], gridDim.x, stg, stw,
                                                     cbs, cb_lds, B.chunk_base + nfp + 1);
}

// Fills n16 16-byte words and then ntail single bytes with zeros (a kernel,
// not hipMemsetAsync: memset nodes captured in HIP graphs have given wrong
// results on replays here, and a caller may capture these launches).
__global__ void k_zero(uint4* p, int64_t n16, unsigned char* tail, int ntail)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n16; k += stride) p[k] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

// verification records of molecules no chunk evaluates: not within the
// cutoff (zero(MinimumDistance), src/minimum_distances.jl:21-28)
__global__ void k_vrec_init(VRec* p, int64_t n)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        VRec o;
        o.within = 0; o.i = -1; o.j = -1; o.d = INFINITY; o.ref = 0; o.pad = 0; o.dref = INFINITY;
        p[k] = o;
    }
}

// integer hits -> FP64 accumulators (assign: the first run of the job, every
// entry written, so the accumulators need not hold zeros)
__global__ void k_flush(ull* cnt, double* acc, ull* hits, int64_t n, int64_t nhist, int64_t half_lo, int64_t half_hi,
                        double w, int assign)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        ull c = cnt[k];
        double mult = (k >= half_lo && k < half_hi) ? w / 2 : w;
        if (assign) {
            acc[k] = (double)c * mult;
        } else {
            if (!c) continue;
            acc[k] += (double)c * mult;
        }
        if (k < nhist) hits[k] += c;
        cnt[k] = 0ull;
    }
}

// Read-back of a job whose frames all had one weight: the md/rdf hits are
// added to the cumulative ones (read back as 64-bit words); each group count
// goes out as a 16-bit word, and the counts that do not fit (0xffff and up)
// are listed as (entry, count) in ov, at most ovcap of them (over[1] counts
// them all).  over[0] is set when a count exceeds limit (at most 2^32 - 1).
// The counts stay in cnt for the FP64 path.
__global__ void k_pack16(const ull* cnt, uint16_t* out, ull* hits, int64_t n, int64_t nhist, ull limit, ull* over,
                         ulonglong2* ov, int64_t ovcap)
{
    bool big = false;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const ull c = cnt[k];
        big |= c > limit;
        if (k < nhist) {
            hits[k] += c;
            continue;
        }
        out[k - nhist] = (uint16_t)(c < 0xffffull ? c : 0xffffull);
        if (c >= 0xffffull) {
            const ull slot = atomicAdd(&over[1], 1ull);
            if (slot < (ull)ovcap) ov[slot] = make_ulonglong2((ull)k, c);
        }
    }
    if (big) atomicOr(&over[0], 1ull);
}

// ---------------------------------------------------------------------------
// many jobs per launch (cm_jobs_submit_device): one round runs one batch of
// frames of every job, and each stage of the single-solute path is one launch
// over all of them.  A device table holds each job's parameters and buffers
// and, per stage, how many blocks of that stage's single-job grid the job has;
// every block of a stage launch walks the concatenated block lists (a block
// index maps to a job by a search over the prefix of the counts), so one
// launch does exactly the work of the jobs' own launches.  The launches read
// nothing but the table's address: a round is a table upload plus the same
// launch sequence, which is captured once in a HIP graph and replayed.
// ---------------------------------------------------------------------------
constexpr int MS_MAX = 128;   // jobs per round
enum MStage {
    MS_ZERO, MS_RADII, MS_GRID, MS_WRAP, MS_SCAN_CELLS, MS_SCATTER, MS_NEAR, MS_DFIELD, MS_PREP, MS_SCAN_ROWS0,
    MS_MOL_SCATTER0, MS_FILL, MS_CHUNK0, MS_BULK_INIT, MS_BULK_COMPACT, MS_PLACE_KEYS, MS_SCAN_ROWSK,
    MS_MOL_SCATTERK, MS_PLACE_NEAR, MS_CHUNKK, MS_GEO_FLAGS, MS_N
};

struct SysDev {
    KP P;
    Bufs B;
    const void* xu;          // device coordinates of the round's frames
    const void* xv;
    uint4* zp[3];            // buffers cleared at the start of the round (cell counts, near flags, key counts)
    int64_t z16[3];          // their sizes in 16-byte words
    int nbx[MS_N];           // x extent (blocks) of each stage's grid for this job; y = blocks / nbx
};

struct MultiTab {
    int nsys;
    int pad;
    int base[MS_N][MS_MAX + 1];   // per stage: prefix over the jobs of their block counts
    SysDev sys[MS_MAX];
};

// Runs fn(job, bx, nbx, by, nby) for every (job, block) of stage ST, the
// block indices of the launch striding over the concatenated lists.  SYNC:
// the stage's body uses shared memory, so the block synchronises between two
// of its virtual blocks.
template <int ST, bool SYNC, class Fn>
__device__ __forceinline__ void multi_blocks(const MultiTab* __restrict__ MT, Fn fn)
{
    // the stage's prefix row in LDS; a block's virtual blocks increase, so
    // its job index only moves forward
    __shared__ int base[MS_MAX + 1];
    const int n = MT->nsys;
    for (int k = threadIdx.x; k <= n; k += blockDim.x) base[k] = MT->base[ST][k];
    __syncthreads();
    // values read from LDS are the same in every lane; made uniform, the
    // job's parameters below are scalar loads, as kernel arguments would be
    const int total = __builtin_amdgcn_readfirstlane(base[n]);
    int lo = 0;
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
        while (base[lo + 1] <= vb) lo++;
        lo = __builtin_amdgcn_readfirstlane(lo);
        const int b0 = __builtin_amdgcn_readfirstlane(base[lo]), b1 = __builtin_amdgcn_readfirstlane(base[lo + 1]);
        const SysDev& S = MT->sys[lo];
        const int local = vb - b0;
        const int nbx = S.nbx[ST];
        const int nby = (b1 - b0) / nbx;
        fn(S, local - (local / nbx) * nbx, nbx, local / nbx, nby);
        if (SYNC) __syncthreads();
    }
}

__global__ void km_zero(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_ZERO, false>(MT, [&](const SysDev& S, int bx, int nbx, int, int) {
        const int64_t n0 = S.z16[0], n1 = S.z16[1], tot = n0 + n1 + S.z16[2];
        for (int64_t k = bx * (int64_t)blockDim.x + threadIdx.x; k < tot; k += (int64_t)nbx * blockDim.x) {
            uint4* q = k < n0 ? S.zp[0] + k : (k < n0 + n1 ? S.zp[1] + (k - n0) : S.zp[2] + (k - n0 - n1));
            *q = make_uint4(0u, 0u, 0u, 0u);
        }
    });
}

template <class T>
__global__ void __launch_bounds__(BLOCK) km_radii(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_RADII, true>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_radii<T>(S.P, S.B, (const T*)S.xu, (const T*)S.xv, 0, bx, nbx, by, nby);
    });
}

__global__ void km_grid(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_GRID, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_grid(S.P, S.B, 0, bx, nbx, by, nby);
    });
}

template <class T>
__global__ void km_wrap_atoms(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_WRAP, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_wrap_atoms<T>(S.P, S.B, (const T*)S.xu, bx, nbx, by, nby);
    });
}

__global__ void __launch_bounds__(1024) km_scan_cells(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_SCAN_CELLS, true>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_scan_cells(S.P, S.B, bx, nbx, by, nby);
    });
}

__global__ void km_scatter(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_SCATTER, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_scatter(S.P, S.B, S.B.nitems, bx, nbx, by, nby);
    });
}

__global__ void km_near_mark(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_NEAR, true>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_near_mark(S.P, S.B, bx, nbx, by, nby);
    });
}

__global__ void km_dfield(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_DFIELD, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_dfield(S.P, S.B, bx, nbx, by, nby);
    });
}

template <class T>
__global__ void km_prep_solvent_tiles(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_PREP, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_prep_solvent_tiles<T>(S.P, S.B, (const T*)S.xv, bx, nbx, by, nby);
    });
}

template <bool RANDOM>
__global__ void __launch_bounds__(1024) km_scan_rows(const MultiTab* __restrict__ MT)
{
    multi_blocks<RANDOM ? MS_SCAN_ROWSK : MS_SCAN_ROWS0, true>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_scan_rows(S.P, S.B, RANDOM ? 1 : 0, RANDOM ? S.P.K : 1, bx, nbx, by, nby);
    });
}

template <bool RANDOM>
__global__ void km_mol_scatter(const MultiTab* __restrict__ MT)
{
    multi_blocks<RANDOM ? MS_MOL_SCATTERK : MS_MOL_SCATTER0, false>(MT, [&](const SysDev& S, int bx, int nbx, int by,
                                                                            int nby) {
        b_mol_scatter(S.P, S.B, RANDOM ? 1 : 0, RANDOM ? S.P.K : 1, bx, nbx, by, nby);
    });
}

template <class T>
__global__ void km_fill_real(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_FILL, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_fill_real<T>(S.P, S.B, (const T*)S.xv, bx, nbx, by, nby);
    });
}

template <bool RANDOM>
__global__ void __launch_bounds__(1024) km_chunk_table(const MultiTab* __restrict__ MT)
{
    multi_blocks<RANDOM ? MS_CHUNKK : MS_CHUNK0, true>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_chunk_table(S.P, S.B, RANDOM ? 1 : 0, RANDOM ? S.P.K : 1, bx, nbx, by, nby);
    });
}

__global__ void km_bulk_init(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_BULK_INIT, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_bulk_init(S.P, S.B, bx, nbx, by, nby);
    });
}

__global__ void __launch_bounds__(1024) km_bulk_compact(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_BULK_COMPACT, true>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_bulk_compact(S.P, S.B, bx, nbx, by, nby);
    });
}

__global__ void km_place_keys(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_PLACE_KEYS, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_place_keys(S.P, S.B, bx, nbx, by, nby);
    });
}

template <class T>
__global__ void km_place_near(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_PLACE_NEAR, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_place_near<T>(S.P, S.B, (const T*)S.xv, bx, nbx, by, nby);
    });
}

__global__ void km_geo_flags(const MultiTab* __restrict__ MT)
{
    multi_blocks<MS_GEO_FLAGS, false>(MT, [&](const SysDev& S, int bx, int nbx, int by, int nby) {
        b_geo_flags(S.P, S.B, bx, nbx, by, nby);
    });
}

// The evaluation over the chunks of every job of the round: chunk g of the
// launch is chunk g - cb[j] of job j, cb the prefix of the jobs' chunk counts
// (read from their chunk tables at the start of the launch; jobs without
// random passes have none in the random launch).
template <class TU, class TV, bool RANDOM, int C, int NVM>
__global__ void __launch_bounds__(64, 3) km_eval_tiles(const MultiTab* __restrict__ MT, int* ctr)
{
    extern __shared__ float4 dyn[];
    __shared__ int cb[MS_MAX + 1];
    const int n = MT->nsys;
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int j = 0; j < n; j++) {
            cb[j] = acc;
            const KP& Pj = MT->sys[j].P;
            const int nfp = Pj.F * (RANDOM ? Pj.K : 1);
            acc += (RANDOM && Pj.K == 0) ? 0 : MT->sys[j].B.chunk_base[nfp];
        }
        cb[n] = acc;
    }
    __syncthreads();
    const int total = __builtin_amdgcn_readfirstlane(cb[n]);
    // a fixed stride, then the last DYN_PCT per cent of the chunks on request
    // from the counter ctr (zero at the start of the launch), as in the
    // single-job evaluation; either way a block's chunks increase
    const int ndyn = (int)((int64_t)total * DYN_PCT / 100);
    const int tdyn = total >= DYN_MIN * (int)gridDim.x ? total - ndyn : total;
    const int dgrab = max(1, ndyn / ((int)gridDim.x * DYN_GRABS));
    int dnext = 0, dend = 0;
    auto next_chunk = [&](int g) {
        if (g < tdyn || tdyn == total) return g;
        if (dnext < dend) return dnext++;
        int v = 0;
        if (threadIdx.x == 0) v = atomicAdd(ctr, dgrab);
        dnext = tdyn + __builtin_amdgcn_readlane(v, 0);
        dend = dnext + dgrab;
        return dnext++;
    };
    int lo = 0;
    for (int g = next_chunk(blockIdx.x); g < total; g = next_chunk(g + gridDim.x)) {
        while (cb[lo + 1] <= g) lo++;
        // uniform (the same in every lane): the job's parameters are scalar loads
        lo = __builtin_amdgcn_readfirstlane(lo);
        const SysDev& S = MT->sys[lo];
        const int ch = g - __builtin_amdgcn_readfirstlane(cb[lo]);
        float4* stg = dyn;
        float* stw = (float*)(dyn + S.P.tcap);
        eval_tiles_chunks<TU, TV, RANDOM, C, NVM, false>(S.P, S.B, (const TU*)S.xu, (const TV*)S.xv, ch, ch + 1, 1, stg,
                                                         stw, nullptr, false, nullptr);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(CM_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// x: blocks over n molecules or slots; y: one per (frame, pass) pair, up to
// the grid limit (the kernels stride over y beyond it)
dim3 grid_slots(int n, int pairs)
{
    return dim3((unsigned)std::max(1, (n + BLOCK - 1) / BLOCK), (unsigned)std::max(1, std::min(pairs, 65535)));
}

int grid_for(int64_t n)
{
    int64_t b = (n + BLOCK - 1) / BLOCK;
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    return (int)b;
}

// zeroes `bytes` bytes at p (16-byte aligned: every buffer here starts on a
// 256-byte boundary) on stream s
int zero_async(void* p, size_t bytes, hipStream_t s)
{
    if (!bytes) return CM_OK;
    const int64_t n16 = (int64_t)(bytes / 16);
    const int ntail = (int)(bytes - 16 * (size_t)n16);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n16 + BLOCK - 1) / BLOCK, 2048))),
                       dim3(BLOCK), 0, s, (uint4*)p, n16, (unsigned char*)p + 16 * n16, ntail);
    HIPCHK(hipGetLastError());
    return CM_OK;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Memory released by finished jobs, kept for the next ones (a process that
// runs one mddf call after another would otherwise pay hipMalloc/hipFree,
// which synchronise the device, for about 1 GB of batch scratch and tens of
// MB of page-locked read-back buffers per call).  Blocks are reused for
// requests of at least half their size; beyond a cap the largest are
// released.  Heap singletons: never destroyed, so nothing runs at exit.
struct BlockCache {
    std::mutex m;
    std::multimap<size_t, std::pair<int, void*>> blocks;   // bytes -> (device, pointer)
    size_t total = 0;
};

BlockCache& dev_cache()
{
    static BlockCache* c = new BlockCache;
    return *c;
}

BlockCache& pin_cache()
{
    static BlockCache* c = new BlockCache;
    return *c;
}

constexpr size_t DEV_CACHE_MAX = (size_t)48 << 30;
constexpr size_t PIN_CACHE_MAX = (size_t)4 << 30;

void* cache_take(BlockCache& c, size_t bytes, int dev, size_t* got)
{
    std::lock_guard<std::mutex> lk(c.m);
    for (auto it = c.blocks.lower_bound(bytes); it != c.blocks.end() && it->first <= 2 * bytes; ++it) {
        if (it->second.first != dev) continue;
        void* p = it->second.second;
        *got = it->first;
        c.total -= it->first;
        c.blocks.erase(it);
        return p;
    }
    return nullptr;
}

// returns a block; blocks beyond the cap (largest first) come back to be freed
std::vector<std::pair<int, void*>> cache_put(BlockCache& c, void* p, size_t bytes, int dev, size_t cap)
{
    std::vector<std::pair<int, void*>> out;
    std::lock_guard<std::mutex> lk(c.m);
    c.blocks.emplace(bytes, std::make_pair(dev, p));
    c.total += bytes;
    while (c.total > cap && !c.blocks.empty()) {
        auto it = std::prev(c.blocks.end());
        c.total -= it->first;
        out.push_back(it->second);
        c.blocks.erase(it);
    }
    return out;
}

std::vector<std::pair<int, void*>> cache_drain(BlockCache& c)
{
    std::lock_guard<std::mutex> lk(c.m);
    std::vector<std::pair<int, void*>> out;
    for (auto& kv : c.blocks) out.push_back(kv.second);
    c.blocks.clear();
    c.total = 0;
    return out;
}

int current_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    return dev;
}

// device block of at least `bytes`: a cached one, else hipMalloc (after
// releasing the cache when the device is full)
int dev_take(DevBuf& b, size_t bytes)
{
    const int dev = current_device();
    size_t got = 0;
    if (void* p = cache_take(dev_cache(), bytes, dev, &got)) {
        b.p = p;
        b.bytes = got;
        return CM_OK;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        for (auto& blk : cache_drain(dev_cache())) (void)hipFree(blk.second);
        e = hipMalloc(&b.p, bytes);
    }
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(CM_ENOMEM, std::string("hipMalloc failed for ") + std::to_string(bytes) + " bytes");
    }
    b.bytes = bytes;
    return CM_OK;
}

// a block no stream uses any more goes back to the cache
void dev_put(DevBuf& b)
{
    if (!b.p) return;
    for (auto& blk : cache_put(dev_cache(), b.p, b.bytes, current_device(), DEV_CACHE_MAX)) (void)hipFree(blk.second);
    b.p = nullptr;
    b.bytes = 0;
}

struct PinBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

int pin_take(PinBuf& b, size_t bytes)
{
    size_t got = 0;
    if (void* p = cache_take(pin_cache(), bytes, 0, &got)) {
        b.p = p;
        b.bytes = got;
        return CM_OK;
    }
    if (hipHostMalloc(&b.p, bytes, hipHostMallocDefault) != hipSuccess) {
        b.p = nullptr;
        return fail(CM_ENOMEM, std::string("hipHostMalloc failed for ") + std::to_string(bytes) + " bytes");
    }
    b.bytes = bytes;
    return CM_OK;
}

void pin_put(PinBuf& b)
{
    if (!b.p) return;
    for (auto& blk : cache_put(pin_cache(), b.p, b.bytes, 0, PIN_CACHE_MAX)) (void)hipHostFree(blk.second);
    b.p = nullptr;
    b.bytes = 0;
}

// Streams of destroyed jobs, kept per device for the next ones: a process
// that runs one mddf call after another would otherwise create and destroy a
// stream per call.  A stream comes back idle (cm_job_destroy waits for it).
struct StreamPool {
    std::mutex m;
    std::vector<std::pair<int, hipStream_t>> s;   // (device, stream)
};

StreamPool& stream_pool()
{
    static StreamPool* p = new StreamPool;
    return *p;
}

constexpr size_t STREAM_POOL_MAX = 64;

int stream_take(hipStream_t* out)
{
    const int dev = current_device();
    {
        StreamPool& p = stream_pool();
        std::lock_guard<std::mutex> lk(p.m);
        for (size_t k = p.s.size(); k-- > 0;)
            if (p.s[k].first == dev) {
                *out = p.s[k].second;
                p.s.erase(p.s.begin() + k);
                return CM_OK;
            }
    }
    if (hipStreamCreateWithFlags(out, hipStreamNonBlocking) != hipSuccess) {
        *out = nullptr;
        return fail(CM_EHIP, "hipStreamCreate failed");
    }
    return CM_OK;
}

void stream_put(hipStream_t st)
{
    if (!st) return;
    int dev = 0;
    if (hipStreamGetDevice(st, &dev) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamDestroy(st);
        return;
    }
    StreamPool& p = stream_pool();
    {
        std::lock_guard<std::mutex> lk(p.m);
        if (p.s.size() < STREAM_POOL_MAX) {
            p.s.emplace_back(dev, st);
            return;
        }
    }
    (void)hipStreamDestroy(st);
}

// host memory the runtime knows as page-locked (hipHostMalloc'd or registered)
bool page_locked(const void* p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // pageable memory is unknown to the runtime
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

void plane_spacings(const double* H, double sp[3])
{
    double V = std::fabs(cm::cell_volume(H));
    for (int k = 0; k < 3; k++) {
        const double *p = H + 3 * ((k + 1) % 3), *q = H + 3 * ((k + 2) % 3);
        double x = p[1] * q[2] - p[2] * q[1], y = p[2] * q[0] - p[0] * q[2], z = p[0] * q[1] - p[1] * q[0];
        sp[k] = V / std::sqrt(x * x + y * y + z * z);
    }
}

}  // namespace

struct cm_job {
    cm_job_desc d;
    KP kp;
    bool atoms_family = true;
    hipStream_t stream = nullptr;
    int64_t Nu = 0, Nv = 0;
    int K = 0, kb = 0;
    // accumulators: [md | md_r | rdf | rdf_r | gu | gv | gu_r | gv_r]
    int64_t acc_len = 0;
    DevBuf cnt, acc, hits, stats;
    double pending_w = 0.0;
    bool pending = false;
    int64_t flushes = 0;   // weight runs already added to the FP64 accumulators
    double volume_total = 0.0;
    DevBuf gmap;   // group maps
    // batch buffers
    DevBuf geo, cells_cnt, cells_start, tmp4, tmpcell, sorted, solw, ucen, vw, vcell, vcen;
    DevBuf P, vw_r, vcen_r, bflag, blist, nbulk, uk, drawn_ptr, drawn_idx;
    DevBuf near, mkey, mrank, mcnt, mstart, msorted, nnear, chunk_base;
    DevBuf xin_u, xin_v;
    double cell_frac = 0.5;   // tiles family: fine cell side as a fraction of the cutoff
    int tile_c = 1;           // tiles family: columns per atom row (1, 2, 4 or 8)
    int eval_grid = 0;        // tiles family: workgroups of the persistent evaluation kernel (0: fill the device)
    int cus = 256;            // compute units of the device
    int Fcap = 0;
    int cap_cells = 0;
    bool finished = false;
    bool profiling = false;
    // profiling: five events per batch (batch start, real-pass evaluation
    // start/end, random-pass evaluation start/end), recorded without a host
    // wait and resolved when the statistics are read
    std::vector<std::array<hipEvent_t, 5>> prof_ev;
    size_t prof_used = 0, prof_done = 0;
    std::vector<char> prof_random;                     // the batch had random passes
    std::vector<std::pair<float, float>> launch_ms;   // per batch: real, random evaluation (ms; -1 when absent)
    hipEvent_t rb_ev[RB_PIECES + 1] = {};   // pipelined read-back (cm_job_finish)
    double eval_ms = 0.0, total_ms = 0.0;
    int64_t eval_launches = 0, total_launches = 0, frames = 0;
    std::vector<Geo> hgeo;
    std::vector<int> huk, hdp, hdi;
    std::vector<int32_t> hgm;        // group maps (host copy of the upload)
    PinBuf hres;                     // accumulators and hit counts read back at finish (page-locked)
    // verification mode: every batch's records, placements and reference
    // draws are copied to the host, keyed by trajectory frame number
    bool verify = false;
    DevBuf vrec;
    struct VerifyFrame {
        std::vector<VRec> rec;            // [NP][Mv] (tiles family)
        std::vector<double> placed;       // [K][Nv][3], NaN where a slot was not placed
        std::vector<uint8_t> placed_flag; // [K][Mv]
        std::vector<int32_t> draws;       // [K]
    };
    std::map<int64_t, VerifyFrame> vstore;
    // buffers carved from a fixed device region instead of allocated (the
    // entry point that allocates nothing); null for ordinary jobs
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
};

namespace {

// Grows a job buffer.  A block being replaced may still be read by work
// queued on the job's stream: the host waits for that stream only (not the
// device, as hipFree would, which would stall the other slots of a request
// server), then the block goes back to the cache for later jobs.  Jobs with
// an arena take the next 256-byte aligned piece of it instead (never given
// back; the arena is reset per call).
int ensure(cm_job* J, DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return CM_OK;
    if (J->arena) {
        const size_t off = (J->arena_used + 255) & ~(size_t)255;
        if (off > J->arena_cap || bytes > J->arena_cap - off)
            return fail(CM_ENOMEM, "job does not fit the fixed device scratch of the no-allocation entry point");
        b.p = J->arena + off;
        b.bytes = bytes;
        J->arena_used = off + bytes;
        return CM_OK;
    }
    if (b.p) {
        if (J->stream) {
            const hipError_t e = hipStreamSynchronize(J->stream);
            if (e != hipSuccess) return fail(CM_EHIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
        }
        dev_put(b);
    }
    b.p = nullptr;
    b.bytes = 0;
    return dev_take(b, bytes);
}

int64_t per_frame_bytes(const cm_job* J, int dtype_size, bool host_input, int64_t cap, int64_t kcap)
{
    int64_t Mu = J->d.nmols_u, Mv = J->d.nmols_v, NP = 1 + J->K;
    int64_t b = sizeof(Geo) + 64;
    if (host_input) b += (J->Nu + J->Nv) * 3 * dtype_size;
    int64_t items = J->atoms_family ? J->d.natoms_per_mol_u : Mu;
    b += items * (16 + 16 + 4) + (cap + 1) * 8;
    if (!J->atoms_family) {
        b += J->Nu * 16 + Mu * 16 + (Mu + 1) * 4 + J->K * 8;
        b += J->Nv * 16 + Mv * 20;
        b += (int64_t)J->K * (J->Nv * (24 + 16) + Mv * 16);
    } else {
        b += cap * 9 + NP * ((kcap + 1) * 8 + Mv * 12 + 4) + 8;
        b += J->Nv * 16 + (int64_t)J->K * J->Nv * (24 + 16);
    }
    b += (int64_t)J->kb * Mv * 5 + J->kb * 4;
    return b;
}

int64_t cells_for(const cm_job* J, const double* cell9, int64_t* keys)
{
    cm::Cell64 c;
    cm::cell_setup(cell9, J->d.triclinic, &c);
    double sp[3];
    plane_spacings(c.H, sp);
    double side = J->atoms_family ? J->kp.cutsys * J->cell_frac : J->kp.cutsys;
    int64_t n = 1, mx = 1;
    for (int k = 0; k < 3; k++) {
        int64_t nk = std::max<int64_t>(1, (int64_t)std::floor(sp[k] / side));
        n *= nk;
        mx = std::max(mx, nk);
    }
    int bits = 0;
    while (((int64_t)1 << bits) < mx) bits++;
    *keys = (int64_t)1 << (3 * bits);
    return n;
}

// grid 0: as many workgroups as the device holds at once (persistent kernel)
template <class TU, class TV, bool RANDOM, int C>
void launch_tiles_one(hipStream_t s, int grid, int cus, const KP& P, const Bufs& B, const void* xu, const void* xv)
{
    const int nfp = P.F * (RANDOM ? P.K : 1);
    const size_t shm = (sizeof(float4) + (P.mol_lane ? sizeof(float) : 0)) * (size_t)P.tcap +
                       sizeof(int) * (size_t)std::min(nfp + 1, CBCAP);
    // 3-atom solvent molecules: one molecule per lane
    // verification (records kept) in its own instantiations, so that the
    // ordinary kernels carry none of its registers
    auto kern = B.vrec ? (P.mol_lane ? k_eval_tiles<TU, TV, RANDOM, C, 3, true> : k_eval_tiles<TU, TV, RANDOM, C, 0, true>)
                       : (P.mol_lane ? k_eval_tiles<TU, TV, RANDOM, C, 3, false> : k_eval_tiles<TU, TV, RANDOM, C, 0, false>);
    if (grid <= 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 64, shm) != hipSuccess || nb < 1) nb = 1;
        grid = nb * cus;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), shm, s, P, B, (const TU*)xu, (const TV*)xv);
}

template <int C>
void launch_tiles(bool f32, bool random, hipStream_t s, int grid, int cus, const KP& P, const Bufs& B, const void* xu,
                  const void* xv)
{
    if (f32) {
        if (random) launch_tiles_one<float, double, true, C>(s, grid, cus, P, B, xu, nullptr);
        else launch_tiles_one<float, float, false, C>(s, grid, cus, P, B, xu, xv);
    } else {
        if (random) launch_tiles_one<double, double, true, C>(s, grid, cus, P, B, xu, nullptr);
        else launch_tiles_one<double, double, false, C>(s, grid, cus, P, B, xu, xv);
    }
}

void launch_tiles_c(int c, bool f32, bool random, hipStream_t s, int grid, int cus, const KP& P, const Bufs& B,
                    const void* xu, const void* xv)
{
    if (c == 8) launch_tiles<8>(f32, random, s, grid, cus, P, B, xu, xv);
    else if (c == 4) launch_tiles<4>(f32, random, s, grid, cus, P, B, xu, xv);
    else if (c == 2) launch_tiles<2>(f32, random, s, grid, cus, P, B, xu, xv);
    else launch_tiles<1>(f32, random, s, grid, cus, P, B, xu, xv);
}

// Verification mode: waits for the batch and keeps, per frame, the records
// of every pass (tiles family), the placed solvent of every random pass (by
// slot) and the drawn reference solute molecules.
int verify_collect(cm_job* J, int F, const int64_t* frame_index)
{
    HIPCHK(hipStreamSynchronize(J->stream));
    const int64_t Mv = J->d.nmols_v, Nv = J->Nv, K = J->K, NP = 1 + K, nv = J->d.natoms_per_mol_v;
    std::vector<VRec> rec;
    if (J->atoms_family) {
        rec.resize((size_t)F * NP * Mv);
        HIPCHK(hipMemcpy(rec.data(), J->vrec.p, sizeof(VRec) * rec.size(), hipMemcpyDeviceToHost));
    }
    std::vector<double> P;
    std::vector<int> msorted, nnear;
    if (K > 0) {
        P.resize((size_t)F * K * Nv * 3);
        HIPCHK(hipMemcpy(P.data(), J->P.p, sizeof(double) * P.size(), hipMemcpyDeviceToHost));
        if (J->atoms_family) {
            msorted.resize((size_t)F * NP * Mv);
            nnear.resize((size_t)F * NP);
            HIPCHK(hipMemcpy(msorted.data(), J->msorted.p, sizeof(int) * msorted.size(), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(nnear.data(), J->nnear.p, sizeof(int) * nnear.size(), hipMemcpyDeviceToHost));
        }
    }
    for (int f = 0; f < F; f++) {
        cm_job::VerifyFrame& vf = J->vstore[frame_index[f]];
        if (J->atoms_family) vf.rec.assign(rec.begin() + (size_t)f * NP * Mv, rec.begin() + (size_t)(f + 1) * NP * Mv);
        vf.placed.assign((size_t)K * Nv * 3, NAN);
        vf.placed_flag.assign((size_t)K * Mv, 0);
        vf.draws.assign((size_t)K, 0);
        for (int64_t k = 0; k < K; k++) {
            const double* src = P.data() + ((size_t)f * K + k) * Nv * 3;
            double* dst = vf.placed.data() + (size_t)k * Nv * 3;
            if (J->atoms_family) {
                // near slots only, stored at their sorted positions
                const int64_t fp = (int64_t)f * NP + 1 + k;
                for (int64_t pos = 0; pos < nnear[fp]; pos++) {
                    const int64_t v = msorted[fp * Mv + pos];
                    std::memcpy(dst + v * nv * 3, src + pos * nv * 3, sizeof(double) * 3 * nv);
                    vf.placed_flag[k * Mv + v] = 1;
                }
            } else {
                std::memcpy(dst, src, sizeof(double) * 3 * Nv);
                std::fill(vf.placed_flag.begin() + k * Mv, vf.placed_flag.begin() + (k + 1) * Mv, (uint8_t)1);
                vf.draws[k] = J->huk[(size_t)f * K + k];
            }
        }
    }
    return CM_OK;
}

// Host geometry of a batch's frames: cells, inverses, plane spacings and, for
// the single-solute kernels, the fine cell grid.
int batch_geo(const cm_job* J, int F, const int64_t* frame_index, const double* cells, int64_t cap, int64_t kcap, Geo* out)
{
    const double cutsys = J->kp.cutsys;
    for (int f = 0; f < F; f++) {
        Geo& g = out[f];
        memset(&g, 0, sizeof(Geo));
        cm::cell_setup(cells + 9 * f, J->d.triclinic, &g.cell);
        for (int k = 0; k < 9; k++) {
            g.H32[k] = (float)g.cell.H[k];
            g.Hi32[k] = (float)g.cell.Hi[k];
        }
        plane_spacings(g.cell.H, g.sp);
        g.ortho = (g.cell.H[1] == 0.0 && g.cell.H[2] == 0.0 && g.cell.H[3] == 0.0 && g.cell.H[5] == 0.0 &&
                   g.cell.H[6] == 0.0 && g.cell.H[7] == 0.0) ? 1 : 0;
        double spmin = std::min(g.sp[0], std::min(g.sp[1], g.sp[2]));
        if (!(cutsys < 0.5 * spmin * (1.0 - 1e-9)))
            return fail(CM_EUNSUPPORTED, "cutoff must be smaller than half of the smallest plane spacing of the unit cell");
        g.frame = (uint32_t)frame_index[f];
        if (J->atoms_family) {
            double side = cutsys * J->cell_frac;
            g.ncell = 1;
            for (int k = 0; k < 3; k++) {
                g.nc[k] = std::max(1, (int)std::floor(g.sp[k] / side));
                g.ncell *= g.nc[k];
            }
            if (g.ncell > cap) return fail(CM_EHIP, "internal error: fine grid larger than its allocation");
            int mx = std::max(g.nc[0], std::max(g.nc[1], g.nc[2]));
            g.kbits = 0;
            while ((1 << g.kbits) < mx) g.kbits++;
            if ((int64_t)1 << (3 * g.kbits) > kcap) return fail(CM_EHIP, "internal error: key space larger than its allocation");
        }
    }
    return CM_OK;
}

// The kernels' view of a job's buffers (geometry array at geo)
// The near flags [F][cap] (bytes) and, after them, the encoded distances of
// the cell centres [F][cap] and lower corners [F][cap] of the tiles family:
// one block, cleared together per batch.
size_t near_flags_bytes(int64_t F, int64_t cap) { return ((size_t)F * cap + 15) & ~(size_t)15; }
size_t near_bytes(int64_t F, int64_t cap) { return near_flags_bytes(F, cap) + (((size_t)8 * F * cap + 15) & ~(size_t)15); }

Bufs batch_bufs(const cm_job* J, Geo* geo)
{
    Bufs B;
    memset(&B, 0, sizeof(B));
    B.geo = geo;
    B.cnt = (int*)J->cells_cnt.p;
    B.start = (int*)J->cells_start.p;
    B.tmp4 = (float4*)J->tmp4.p;
    B.tmpcell = (int*)J->tmpcell.p;
    B.sorted = (float4*)J->sorted.p;
    B.nitems = J->atoms_family ? J->kp.nu : J->kp.Mu;
    B.solw = (float4*)J->solw.p;
    B.ucen = (float4*)J->ucen.p;
    B.vw = (float4*)J->vw.p;
    B.vcell = (int*)J->vcell.p;
    B.vcen = (float4*)J->vcen.p;
    B.P = (double*)J->P.p;
    B.vw_r = (float4*)J->vw_r.p;
    B.vcen_r = (float4*)J->vcen_r.p;
    B.bflag = (unsigned char*)J->bflag.p;
    B.blist = (int*)J->blist.p;
    B.nbulk = (int*)J->nbulk.p;
    B.uk = (int*)J->uk.p;
    B.drawn_ptr = (int*)J->drawn_ptr.p;
    B.drawn_idx = (int*)J->drawn_idx.p;
    B.near = (unsigned char*)J->near.p;
    B.mkey = (int*)J->mkey.p;
    B.mrank = (int*)J->mrank.p;
    B.mcnt = (int*)J->mcnt.p;
    B.mstart = (int*)J->mstart.p;
    B.msorted = (int*)J->msorted.p;
    B.nnear = (int*)J->nnear.p;
    B.chunk_base = (int*)J->chunk_base.p;
    ull* c = (ull*)J->cnt.p;
    const int64_t nb = J->d.nbins, ngu = J->d.n_groups_u, ngv = J->d.n_groups_v;
    B.cnt_md = c; B.cnt_md_r = c + nb; B.cnt_rdf = c + 2 * nb; B.cnt_rdf_r = c + 3 * nb;
    B.cnt_gu = c + 4 * nb; B.cnt_gv = B.cnt_gu + ngu * nb; B.cnt_gu_r = B.cnt_gv + ngv * nb; B.cnt_gv_r = B.cnt_gu_r + ngu * nb;
    B.stats = (ull*)J->stats.p;
    return B;
}

int run_batch(cm_job* J, int F, const int64_t* frame_index, const void* xu_dev, const void* xv_dev, int dtype,
              const double* cells, int64_t cap, int64_t kcap)
{
    KP P = J->kp;
    P.F = F;
    P.cap = (int)cap;
    P.kcap = (int)kcap;
    const bool tiles = J->atoms_family;
    // geometry on the host: cells, inverse, spacings (and the fine grid of the tiles family)
    J->hgeo.assign(F, Geo());
    if (int r = batch_geo(J, F, frame_index, cells, cap, kcap, J->hgeo.data())) return r;
    hipStream_t s = J->stream;
    HIPCHK(hipMemcpyAsync(J->geo.p, J->hgeo.data(), sizeof(Geo) * F, hipMemcpyHostToDevice, s));
    Bufs B = batch_bufs(J, (Geo*)J->geo.p);
    B.dfield = (unsigned*)((char*)J->near.p + near_flags_bytes(F, cap));
    B.vrec = (J->verify && tiles) ? (VRec*)J->vrec.p : nullptr;
    if (B.vrec) {
        const int64_t nrec = (int64_t)F * P.NP * P.Mv;
        hipLaunchKernelGGL(k_vrec_init, dim3(grid_for(nrec)), dim3(BLOCK), 0, s, B.vrec, nrec);
    }

    // random draws of the reference solute molecules (mddf_frame!, src/mddf.jl:374-376);
    // with one solute molecule every draw is molecule 0
    if (!tiles && J->K > 0) {
        J->huk.assign((size_t)F * J->K, 0);
        J->hdp.assign((size_t)F * (P.Mu + 1), 0);
        J->hdi.assign((size_t)F * J->K, 0);
        for (int f = 0; f < F; f++) {
            int* dp = &J->hdp[(size_t)f * (P.Mu + 1)];
            for (int k = 0; k < J->K; k++) {
                cm::u32x4 r = cm::philox4x32_10({(uint32_t)k, 0u, (uint32_t)frame_index[f], cm::TAG_REF}, P.key0, P.key1);
                int u = (int)cm::draw_index(r.x, P.Mu);
                J->huk[(size_t)f * J->K + k] = u;
                dp[u + 1]++;
            }
            for (int u = 0; u < P.Mu; u++) dp[u + 1] += dp[u];
            std::vector<int> fill(P.Mu, 0);
            for (int k = 0; k < J->K; k++) {
                int u = J->huk[(size_t)f * J->K + k];
                J->hdi[(size_t)f * J->K + dp[u] + fill[u]++] = k;
            }
        }
        HIPCHK(hipMemcpyAsync(J->uk.p, J->huk.data(), sizeof(int) * J->huk.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(J->drawn_ptr.p, J->hdp.data(), sizeof(int) * J->hdp.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(J->drawn_idx.p, J->hdi.data(), sizeof(int) * J->hdi.size(), hipMemcpyHostToDevice, s));
    } else if (J->K > 0) {
        if (int r = zero_async(J->uk.p, sizeof(int) * (size_t)F * J->K, s)) return r;
    }
    if (int r = zero_async(J->cells_cnt.p, sizeof(int) * (size_t)F * (cap + 1), s)) return r;
    if (tiles) {
        if (int r = zero_async(J->near.p, near_bytes(F, cap), s)) return r;
        if (int r = zero_async(J->mcnt.p, sizeof(int) * (size_t)F * P.NP * (kcap + 1), s)) return r;
    }

    const bool f32 = dtype == CM_F32;
    const float* uf = (const float*)xu_dev;
    const float* vf = (const float*)xv_dev;
    const double* ud = (const double*)xu_dev;
    const double* vd = (const double*)xv_dev;
    const size_t shm = sizeof(unsigned) * 2 * P.nbins;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    if (J->profiling) {
        if (J->prof_used == J->prof_ev.size()) {
            std::array<hipEvent_t, 5> a{};
            for (auto& e : a) HIPCHK(hipEventCreate(&e));
            J->prof_ev.push_back(a);
        }
        for (int k = 0; k < 5; k++) ev[k] = J->prof_ev[J->prof_used][k];
        J->prof_used++;
        HIPCHK(hipEventRecord(ev[4], s));
        J->prof_random.push_back(P.K > 0 ? 1 : 0);
    }
    const int mols = tiles ? 0 : 1;
    {
        int64_t n = (int64_t)F * (P.Mv + (mols ? P.Mu : 0));
        {
            const int64_t nm = (int64_t)P.Mv + (mols ? P.Mu : 0);
            const dim3 gr((unsigned)std::min<int64_t>((nm + BLOCK - 1) / BLOCK, 64), (unsigned)F);
            if (f32) hipLaunchKernelGGL(k_radii<float>, gr, dim3(BLOCK), 0, s, P, B, uf, vf, mols);
            else hipLaunchKernelGGL(k_radii<double>, gr, dim3(BLOCK), 0, s, P, B, ud, vd, mols);
        }
    }
    hipLaunchKernelGGL(k_grid, dim3(1), dim3(64), 0, s, P, B, mols);
    if (tiles) {
        int64_t n = (int64_t)F * P.nu;
        if (f32) hipLaunchKernelGGL(k_wrap_atoms<float>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, uf);
        else hipLaunchKernelGGL(k_wrap_atoms<double>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, ud);
    } else {
        int64_t n = (int64_t)F * P.Mu;
        if (f32) hipLaunchKernelGGL(k_prep_solute_mols<float>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, uf);
        else hipLaunchKernelGGL(k_prep_solute_mols<double>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, ud);
    }
    hipLaunchKernelGGL(k_scan_cells, dim3(F), dim3(1024), 0, s, P, B);
    hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)F * B.nitems)), dim3(BLOCK), 0, s, P, B, B.nitems);
    if (tiles) {
        hipLaunchKernelGGL(k_near_mark, dim3((P.nu + BLOCK - 1) / BLOCK, std::min(F, 65535)), dim3(BLOCK), 0, s, P, B);
        hipLaunchKernelGGL(k_dfield, dim3(grid_for((int64_t)F * cap * 16)), dim3(BLOCK), 0, s, P, B);
        int64_t n = (int64_t)F * P.Mv;
        if (f32) hipLaunchKernelGGL(k_prep_solvent_tiles<float>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, vf);
        else hipLaunchKernelGGL(k_prep_solvent_tiles<double>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, vd);
        hipLaunchKernelGGL(k_scan_rows, dim3(F), dim3(1024), 0, s, P, B, 0, 1);
        hipLaunchKernelGGL(k_mol_scatter, grid_slots(P.Mv, F), dim3(BLOCK), 0, s, P, B, 0, 1);
        if (f32) hipLaunchKernelGGL(k_fill_real<float>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, vf);
        else hipLaunchKernelGGL(k_fill_real<double>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, vd);
        hipLaunchKernelGGL(k_chunk_table, dim3(1), dim3(1024), 0, s, P, B, 0, 1);
    } else {
        int64_t n = (int64_t)F * P.Mv;
        if (f32) hipLaunchKernelGGL(k_prep_solvent<float>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, vf);
        else hipLaunchKernelGGL(k_prep_solvent<double>, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, B, vd);
    }
    if (P.K > 0) hipLaunchKernelGGL(k_bulk_init, dim3(grid_for((int64_t)F * P.kb * P.Mv)), dim3(BLOCK), 0, s, P, B);
    HIPCHK(hipGetLastError());
    // real pass
    if (J->profiling) HIPCHK(hipEventRecord(ev[0], s));
    if (tiles) {
        launch_tiles_c(J->tile_c, f32, false, s, J->eval_grid, J->cus, P, B, xu_dev, xv_dev);
    } else {
        int64_t n = (int64_t)F * P.Mv;
        const int gwords = P.glds ? P.nbins * (P.gnu + (P.self_ ? 0 : P.gnv)) : 0;
        const size_t shq = sizeof(unsigned) * ((2 * P.nbins + gwords + 3) & ~3) + sizeof(int4) * QCAP * (BLOCK / 64) +
                           (P.self_ ? sizeof(int4) * 2 * DCAP * (BLOCK / 64) : 0);
        // small systems: S lanes per solvent molecule split its neighbour walk
        // so that the launch still has about 6 wavefronts per SIMD of work
        const int64_t lanes = (int64_t)P.Mv * F;
        const int S = (int)std::min<int64_t>(14, std::max<int64_t>(1, (int64_t)(J->cus * 4 * 6 * 64) / std::max<int64_t>(lanes, 1)));
        const int gx = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)P.Mv * S + BLOCK - 1) / BLOCK, 4096 / F + 1));
        const dim3 gr((unsigned)(8 * ((gx * F + 7) / 8)));
        // self-MDDF of 3- and 8-atom molecules (water, urea-like) keeps both
        // molecules of a pair in registers
        const int nfix = (P.self_ && P.nu == P.nv && (P.nu == 3 || P.nu == 8)) ? P.nu : 0;
        if (f32) {
            if (nfix == 3) hipLaunchKernelGGL((k_eval_mols_real<float, float, 3>), gr, dim3(BLOCK), shq, s, P, B, uf, vf, gx, S);
            else if (nfix == 8) hipLaunchKernelGGL((k_eval_mols_real<float, float, 8>), gr, dim3(BLOCK), shq, s, P, B, uf, vf, gx, S);
            else hipLaunchKernelGGL((k_eval_mols_real<float, float, 0>), gr, dim3(BLOCK), shq, s, P, B, uf, vf, gx, S);
        } else {
            if (nfix == 3) hipLaunchKernelGGL((k_eval_mols_real<double, double, 3>), gr, dim3(BLOCK), shq, s, P, B, ud, vd, gx, S);
            else if (nfix == 8) hipLaunchKernelGGL((k_eval_mols_real<double, double, 8>), gr, dim3(BLOCK), shq, s, P, B, ud, vd, gx, S);
            else hipLaunchKernelGGL((k_eval_mols_real<double, double, 0>), gr, dim3(BLOCK), shq, s, P, B, ud, vd, gx, S);
        }
    }
    if (J->profiling) HIPCHK(hipEventRecord(ev[1], s));
    HIPCHK(hipGetLastError());
    // random passes
    if (P.K > 0) {
        hipLaunchKernelGGL(k_bulk_compact, dim3(F * P.kb), dim3(1024), 0, s, P, B);
        if (tiles) {
            hipLaunchKernelGGL(k_place_keys, grid_slots(P.Mv, F * P.K), dim3(BLOCK), 0, s, P, B);
            hipLaunchKernelGGL(k_scan_rows, dim3(F * P.K), dim3(1024), 0, s, P, B, 1, P.K);
            hipLaunchKernelGGL(k_mol_scatter, grid_slots(P.Mv, F * P.K), dim3(BLOCK), 0, s, P, B, 1, P.K);
            if (f32) hipLaunchKernelGGL(k_place_near<float>, grid_slots(P.Mv, F * P.K), dim3(BLOCK), 0, s, P, B, vf);
            else hipLaunchKernelGGL(k_place_near<double>, grid_slots(P.Mv, F * P.K), dim3(BLOCK), 0, s, P, B, vd);
            hipLaunchKernelGGL(k_chunk_table, dim3(1), dim3(1024), 0, s, P, B, 1, P.K);
        } else {
            if (f32) hipLaunchKernelGGL(k_place<float>, grid_slots(P.Mv, F * P.K), dim3(BLOCK), 0, s, P, B, vf);
            else hipLaunchKernelGGL(k_place<double>, grid_slots(P.Mv, F * P.K), dim3(BLOCK), 0, s, P, B, vd);
        }
        HIPCHK(hipGetLastError());
        if (J->profiling) HIPCHK(hipEventRecord(ev[2], s));
        if (tiles) {
            launch_tiles_c(J->tile_c, f32, true, s, J->eval_grid, J->cus, P, B, xu_dev, nullptr);
        } else {
            const int gwords = P.glds ? P.nbins * (P.gnu + (P.self_ ? 0 : P.gnv)) : 0;
            const size_t shr = sizeof(unsigned) * (2 * P.nbins + gwords);
            const int64_t n = (int64_t)F * P.K * P.Mv;
            if (f32) hipLaunchKernelGGL(k_eval_mols_rand<float>, dim3(grid_for(n)), dim3(BLOCK), shr, s, P, B, uf);
            else hipLaunchKernelGGL(k_eval_mols_rand<double>, dim3(grid_for(n)), dim3(BLOCK), shr, s, P, B, ud);
        }
        if (J->profiling) HIPCHK(hipEventRecord(ev[3], s));
    }
    // geometry errors detected on the device: kept in the job's error bits and
    // reported when the job is collected, so the host does not wait for the
    // batch here and prepares the next one while it runs
    hipLaunchKernelGGL(k_geo_flags, dim3(1), dim3(64), 0, s, P, B);
    HIPCHK(hipGetLastError());
    J->frames += F;
    if (J->verify) return verify_collect(J, F, frame_index);
    return CM_OK;
}

// sizes of the buffers cleared per batch, rounded up to whole 16-byte words
// (the multi-job path clears them with 16-byte stores)
size_t round16(size_t b) { return (b + 15) & ~(size_t)15; }

int alloc_batch(cm_job* J, int F, int64_t cap, int64_t kcap, int dtype_size, bool host_input)
{
    int r;
    int64_t Mu = J->d.nmols_u, Mv = J->d.nmols_v, NP = 1 + J->K;
    int64_t items = J->atoms_family ? J->d.natoms_per_mol_u : Mu;
    if ((r = ensure(J, J->geo, sizeof(Geo) * F))) return r;
    if ((r = ensure(J, J->cells_cnt, round16(sizeof(int) * F * (cap + 1))))) return r;
    if ((r = ensure(J, J->cells_start, sizeof(int) * F * (cap + 1)))) return r;
    if ((r = ensure(J, J->tmp4, 16 * F * items))) return r;
    if ((r = ensure(J, J->tmpcell, 4 * F * items))) return r;
    if ((r = ensure(J, J->sorted, 16 * F * items))) return r;
    if ((r = ensure(J, J->vw, 16 * F * J->Nv))) return r;
    if (!J->atoms_family) {
        if ((r = ensure(J, J->solw, 16 * F * J->Nu))) return r;
        if ((r = ensure(J, J->ucen, 16 * F * Mu))) return r;
        if ((r = ensure(J, J->drawn_ptr, 4 * F * (Mu + 1)))) return r;
        if ((r = ensure(J, J->drawn_idx, 4 * F * std::max(J->K, 1)))) return r;
        if ((r = ensure(J, J->vcell, 4 * F * Mv))) return r;
        if ((r = ensure(J, J->vcen, 16 * F * Mv))) return r;
    } else {
        if ((r = ensure(J, J->near, near_bytes(F, cap)))) return r;
        if ((r = ensure(J, J->mkey, (size_t)4 * F * NP * Mv))) return r;
        if ((r = ensure(J, J->mrank, (size_t)4 * F * NP * Mv))) return r;
        if ((r = ensure(J, J->mcnt, round16((size_t)4 * F * NP * (kcap + 1))))) return r;
        if ((r = ensure(J, J->mstart, (size_t)4 * F * NP * (kcap + 1)))) return r;
        if ((r = ensure(J, J->msorted, (size_t)4 * F * NP * Mv))) return r;
        if ((r = ensure(J, J->nnear, (size_t)4 * F * NP))) return r;
        if ((r = ensure(J, J->chunk_base, (size_t)4 * (F * std::max(J->K, 1) + 2)))) return r;
    }
    if (J->K > 0) {
        if ((r = ensure(J, J->P, (size_t)24 * F * J->K * J->Nv))) return r;
        if ((r = ensure(J, J->vw_r, (size_t)16 * F * J->K * J->Nv))) return r;
        if (!J->atoms_family) {
            if ((r = ensure(J, J->vcen_r, (size_t)16 * F * J->K * Mv))) return r;
        }
        if ((r = ensure(J, J->bflag, (size_t)F * J->kb * Mv))) return r;
        if ((r = ensure(J, J->blist, (size_t)4 * F * J->kb * Mv))) return r;
        if ((r = ensure(J, J->nbulk, (size_t)4 * F * J->kb))) return r;
        if ((r = ensure(J, J->uk, (size_t)4 * F * J->K))) return r;
    }
    if (J->verify && J->atoms_family)
        if ((r = ensure(J, J->vrec, sizeof(VRec) * (size_t)F * NP * Mv))) return r;
    if (host_input) {
        if ((r = ensure(J, J->xin_u, (size_t)dtype_size * F * J->Nu * 3))) return r;
        if ((r = ensure(J, J->xin_v, (size_t)dtype_size * F * J->Nv * 3))) return r;
    }
    return CM_OK;
}

// the self-MDDF half range of the accumulators: both group hits of a record
// count w/2 (update_counters.jl:48-53)
void half_range(const cm_job* J, int64_t* lo, int64_t* hi)
{
    *lo = *hi = 0;
    if (J->d.autocorrelation) {
        *lo = 4 * (int64_t)J->d.nbins;
        *hi = J->acc_len;
    }
}

// assign: the hits of the job's first weight run go into accumulators that
// may hold anything (the packed read-back reuses their storage); with
// count_hits == false the md/rdf hits are not added to the cumulative ones
// (they already were)
int flush(cm_job* J, double w, hipStream_t stream = nullptr, bool assign = false, bool count_hits = true)
{
    int64_t nb = J->d.nbins;
    int64_t half_lo, half_hi;
    half_range(J, &half_lo, &half_hi);
    hipLaunchKernelGGL(k_flush, dim3(grid_for(J->acc_len)), dim3(BLOCK), 0, stream ? stream : J->stream, (ull*)J->cnt.p,
                       (double*)J->acc.p,
                       (ull*)J->hits.p, J->acc_len, count_hits ? 4 * nb : 0, half_lo, half_hi, w, assign ? 1 : 0);
    HIPCHK(hipGetLastError());
    J->flushes++;
    return CM_OK;
}

// cell slots per frame (the largest grid of the frames, with 10% headroom so
// that batches of later frames of a fluctuating box reuse the allocation) and
// frames per batch
int batch_plan(const cm_job* J, int32_t nframes, const double* cells, int ds, bool host_input, int64_t* capmax_,
               int64_t* kcapmax_, int64_t* Fmax_, int64_t* pfb_)
{
    int64_t capmax = 1, kcapmax = 1;
    for (int f = 0; f < nframes; f++) {
        int64_t kk;
        capmax = std::max(capmax, cells_for(J, cells + 9 * f, &kk));
        kcapmax = std::max(kcapmax, kk);
    }
    if (capmax > (int64_t)1 << 26 || kcapmax > (int64_t)1 << 27)
        return fail(CM_EUNSUPPORTED, "unit cell too large for the cell grid");
    capmax = capmax + capmax / 10 + 1;
    if (!J->atoms_family) kcapmax = 1;
    const int64_t pfb = per_frame_bytes(J, ds, host_input, capmax, kcapmax);
    const int64_t budget = J->d.memory_budget > 0 ? J->d.memory_budget : (int64_t)4 << 30;
    int64_t Fmax = std::max<int64_t>(1, budget / std::max<int64_t>(pfb, 1));
    if (J->d.max_batch_frames > 0) Fmax = std::min<int64_t>(Fmax, J->d.max_batch_frames);
    // at most 48 frames per batch: the random-pass evaluation of config-2 sized
    // jobs ran 25% slower per frame with 64-frame batches than with 32-48
    Fmax = std::min<int64_t>(Fmax, 48);
    *capmax_ = capmax;
    *kcapmax_ = kcapmax;
    *Fmax_ = Fmax;
    *pfb_ = pfb;
    return CM_OK;
}

int submit_common(cm_job* J, int32_t nframes, const int64_t* frame_index, const void* xu, const void* xv, int32_t dtype,
                  const double* cells, const double* weights, bool host_input)
{
    if (!J) return fail(CM_EINVAL, "null job");
    if (J->finished) return fail(CM_ESTATE, "job already finished");
    if (nframes < 0 || (nframes > 0 && (!frame_index || !xu || !xv || !cells || !weights)))
        return fail(CM_EINVAL, "invalid frame arrays");
    if (dtype != CM_F64 && dtype != CM_F32) return fail(CM_EINVAL, "dtype must be CM_F64 or CM_F32");
    int ds = dtype == CM_F32 ? 4 : 8;
    for (int f = 0; f < nframes; f++) {
        if (!(weights[f] > 0.0) || !std::isfinite(weights[f])) return fail(CM_EINVAL, "frame weights must be positive and finite");
        if (frame_index[f] < 1 || frame_index[f] > 0xFFFFFFFFll) return fail(CM_EINVAL, "frame_index must be in [1, 2^32)");
    }
    if (nframes == 0) return CM_OK;
    int64_t capmax, kcapmax, Fmax, pfb;
    int r0 = batch_plan(J, nframes, cells, ds, host_input, &capmax, &kcapmax, &Fmax, &pfb);
    if (r0) return r0;
    int f0 = 0;
    while (f0 < nframes) {
        // a batch holds frames of one weight (the weight multiplies the integer hits at flush)
        double w = weights[f0];
        // the frames of this weight in equal batches (200 frames: 5 x 40, not
        // 4 x 48 + 8)
        int run = 1;
        while (f0 + run < nframes && weights[f0 + run] == w) run++;
        const int nb = (int)((run + Fmax - 1) / Fmax);
        const int F = (run + nb - 1) / nb;
        if (J->pending && J->pending_w != w) {
            int r = flush(J, J->pending_w);
            if (r) return r;
        }
        J->pending = true;
        J->pending_w = w;
        const int64_t cap = capmax;
        for (int f = 0; f < F; f++) {
            cm::Cell64 c;
            cm::cell_setup(cells + 9 * (f0 + f), J->d.triclinic, &c);
            J->volume_total += w * cm::cell_volume(c.H);
        }
        int r = alloc_batch(J, F, cap, kcapmax, ds, host_input);
        if (r) return r;
        const void* du;
        const void* dv;
        if (host_input) {
            const void* su = (const char*)xu + (size_t)f0 * J->Nu * 3 * ds;
            const void* sv = (const char*)xv + (size_t)f0 * J->Nv * 3 * ds;
            HIPCHK(hipMemcpyAsync(J->xin_u.p, su, (size_t)F * J->Nu * 3 * ds, hipMemcpyHostToDevice, J->stream));
            HIPCHK(hipMemcpyAsync(J->xin_v.p, sv, (size_t)F * J->Nv * 3 * ds, hipMemcpyHostToDevice, J->stream));
            // Copies from pageable memory are staged before hipMemcpyAsync
            // returns; copies from page-locked memory run later, in stream
            // order.  The caller may reuse its arrays once the call returns,
            // so page-locked sources are waited for.
            if (page_locked(su) || page_locked(sv)) HIPCHK(hipStreamSynchronize(J->stream));
            du = J->xin_u.p;
            dv = J->xin_v.p;
        } else {
            du = (const char*)xu + (size_t)f0 * J->Nu * 3 * ds;
            dv = (const char*)xv + (size_t)f0 * J->Nv * 3 * ds;
        }
        r = run_batch(J, F, frame_index + f0, du, dv, dtype, cells + 9 * f0, cap, kcapmax);
        if (r) return r;
        f0 += F;
    }
    return CM_OK;
}

// argument checks of a job description (the messages the host maps to
// ArgumentError)
int validate_desc(const cm_job_desc& d)
{
    if (d.nmols_u < 1 || d.natoms_per_mol_u < 1 || d.nmols_v < 1 || d.natoms_per_mol_v < 1)
        return fail(CM_EINVAL, "selections must hold at least one molecule of at least one atom");
    if (d.irefatom < 1 || d.irefatom > d.natoms_per_mol_v)
        return fail(CM_EINVAL, "irefatom must be between 1 and natoms_per_mol_v");
    if (d.nbins < 1 || d.nbins > MAX_NBINS) return fail(CM_EINVAL, "nbins out of range [1, 8192]");
    if (!(d.binstep > 0.0) || !(d.cutoff > 0.0) || !(d.dbulk > 0.0)) return fail(CM_EINVAL, "binstep, cutoff and dbulk must be positive");
    if ((int64_t)cm::setbin(d.cutoff, d.binstep) != d.nbins) return fail(CM_EINVAL, "nbins must equal setbin(cutoff, binstep)");
    if (d.usecutoff ? !(d.dbulk < d.cutoff) : !(d.dbulk <= d.cutoff))
        return fail(CM_EINVAL, "dbulk must not exceed cutoff (the histogram range)");
    if (d.n_random_samples < 1 && !d.coordination_number_only) return fail(CM_EINVAL, "n_random_samples must be greater than 0");
    if (d.n_groups_u < 1 || d.n_groups_v < 1) return fail(CM_EINVAL, "group counts must be positive");
    if (!d.groups_u_ptr && d.n_groups_u != d.natoms_per_mol_u) return fail(CM_EINVAL, "n_groups_u must equal natoms_per_mol_u without custom groups");
    if (!d.groups_v_ptr && d.n_groups_v != (d.autocorrelation ? d.natoms_per_mol_u : d.natoms_per_mol_v))
        return fail(CM_EINVAL, "n_groups_v must equal the atoms per molecule without custom groups");
    if (d.autocorrelation && (d.nmols_u != d.nmols_v || d.natoms_per_mol_u != d.natoms_per_mol_v))
        return fail(CM_EINVAL, "autocorrelation needs identical solute and solvent selections");
    if (d.seed == 0) return fail(CM_EINVAL, "seed must be positive");
    int64_t Nu = (int64_t)d.nmols_u * d.natoms_per_mol_u, Nv = (int64_t)d.nmols_v * d.natoms_per_mol_v;
    if (Nu > INT32_MAX / 4 || Nv > INT32_MAX / 4) return fail(CM_EINVAL, "selections too large");
    if (d.groups_u_ptr) {
        if (!d.groups_u_idx) return fail(CM_EINVAL, "groups_u_idx missing");
        for (int64_t k = 0; k < d.groups_u_ptr[d.natoms_per_mol_u]; k++)
            if (d.groups_u_idx[k] < 0 || d.groups_u_idx[k] >= d.n_groups_u) return fail(CM_EINVAL, "solute group id out of range");
    }
    if (d.groups_v_ptr) {
        if (!d.groups_v_idx) return fail(CM_EINVAL, "groups_v_idx missing");
        const int64_t lim = d.autocorrelation ? d.n_groups_u : d.n_groups_v;
        for (int64_t k = 0; k < d.groups_v_ptr[Nv]; k++)
            if (d.groups_v_idx[k] < 0 || d.groups_v_idx[k] >= lim) return fail(CM_EINVAL, "solvent group id out of range");
    }
    return CM_OK;
}

// (re)initialises a job on its stream for a validated description; device
// buffers already held by J are kept and grown as needed
// host-side fields of a job for a validated description (no device work)
void job_params(cm_job* J, const cm_job_desc& d)
{
    const int64_t Nu = (int64_t)d.nmols_u * d.natoms_per_mol_u, Nv = (int64_t)d.nmols_v * d.natoms_per_mol_v;
    J->d = d;
    J->Nu = Nu;
    J->Nv = Nv;
    J->pending = false;
    J->pending_w = 0.0;
    J->flushes = 0;
    J->volume_total = 0.0;
    J->finished = false;
    J->profiling = false;
    J->eval_ms = J->total_ms = 0.0;
    J->eval_launches = J->total_launches = J->frames = 0;
    J->prof_used = J->prof_done = 0;
    J->verify = false;
    J->vstore.clear();
    J->prof_random.clear();
    J->launch_ms.clear();
    J->cell_frac = 0.5;
    J->eval_grid = 0;
    // tiles family: one solute molecule, solvent molecules that fit a wavefront
    J->atoms_family = (d.nmols_u == 1 && !d.autocorrelation && d.natoms_per_mol_v <= 64);
    J->K = d.coordination_number_only ? 0 : d.n_random_samples;
    J->kb = J->K > 0 ? (J->atoms_family ? 1 : J->K) : 0;
    KP& P = J->kp;
    memset(&P, 0, sizeof(P));
    P.Mu = d.nmols_u; P.nu = d.natoms_per_mol_u; P.Mv = d.nmols_v; P.nv = d.natoms_per_mol_v;
    P.self_ = d.autocorrelation ? 1 : 0;
    P.iref = d.irefatom - 1;
    P.nbins = d.nbins;
    P.usecutoff = d.usecutoff ? 1 : 0;
    P.binstep = d.binstep;
    P.cutsys = d.usecutoff ? d.cutoff : d.dbulk;
    P.cut2 = P.cutsys * P.cutsys;
    P.dbulk = d.dbulk;
    P.inv_binstep = (float)(1.0 / d.binstep);
    {
        auto down = [](double x) { float f = (float)x; return (double)f > x ? std::nextafter(f, -INFINITY) : f; };
        auto up = [](double x) { float f = (float)x; return (double)f < x ? std::nextafter(f, INFINITY) : f; };
        const double bs2 = d.binstep * d.binstep, db2 = d.dbulk * d.dbulk;
        P.c2lo32 = down(P.cut2);
        P.c2hi32 = up(P.cut2);
        P.bs2m32 = down(bs2 * (1.0 - 4e-7));
        P.bs2p32 = up(bs2 * (1.0 + 4e-7));
        P.db2lo32 = down(db2 * (1.0 - 4e-7));
        P.db2hi32 = up(db2 * (1.0 + 4e-7));
    }
    P.K = J->K;
    P.kb = J->kb;
    P.key0 = (uint32_t)d.seed;
    P.key1 = (uint32_t)(d.seed >> 32);
    P.Nu = Nu;
    P.Nv = Nv;
    P.NP = 1 + J->K;
    // 3-atom solvent molecules: one molecule per lane row, 4 columns (16
    // molecules per chunk); otherwise one column, every lane a solvent atom
    // (the per-chunk costs of smaller chunks outweigh their shorter staged
    // lists)
    P.mol_lane = P.nv == 3 ? 1 : 0;
    J->tile_c = P.mol_lane ? 4 : 1;
    P.mpt = P.mol_lane ? 64 / J->tile_c : 64 / P.nv;
    // molecule per lane stages |s'|^2 beside each atom: fewer per round so
    // that 12 one-wavefront blocks per CU still fit
    P.tcap = P.mol_lane ? TCAP_MOL : TCAP_DEFAULT;
    P.gnu = d.n_groups_u;
    P.gnv = d.n_groups_v;
    // molecule-pair kernels keep the group counts in LDS when they fit
    P.glds = (int64_t)d.nbins * (d.n_groups_u + (d.autocorrelation ? 0 : d.n_groups_v)) * 4 <= GLDS_MAX ? 1 : 0;
    {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess)
            if (hipDeviceGetAttribute(&J->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || J->cus < 1)
                J->cus = 256;
    }
    J->acc_len = 4 * (int64_t)d.nbins + 2 * (int64_t)d.n_groups_u * d.nbins + 2 * (int64_t)d.n_groups_v * d.nbins;
}

int job_setup(cm_job* J, const cm_job_desc& d)
{
    job_params(J, d);
    KP& P = J->kp;
    const int64_t Nv = J->Nv;
    const int64_t nb = d.nbins;
    int r;
    if ((r = ensure(J, J->cnt, 8 * J->acc_len))) return r;
    if ((r = ensure(J, J->acc, 8 * J->acc_len))) return r;
    if ((r = ensure(J, J->hits, 8 * 4 * nb))) return r;
    if ((r = ensure(J, J->stats, 64))) return r;
    if ((r = zero_async(J->cnt.p, 8 * J->acc_len, J->stream))) return r;
    if ((r = zero_async(J->acc.p, 8 * J->acc_len, J->stream))) return r;
    if ((r = zero_async(J->hits.p, 8 * 4 * nb, J->stream))) return r;
    if ((r = zero_async(J->stats.p, 64, J->stream))) return r;
    // The page-locked read-back buffer is taken now, while the device is idle
    // for this job: page-locking tens of MB (hipHostMalloc) while the job's
    // kernels run stalled one of them about fourfold (a 2.4 ms evaluation took
    // 9.6 ms in the first job of a process, where the cache has no block yet).
    if (!J->arena) {
        const size_t need = 8 * (size_t)J->acc_len + 8 * 4 * (size_t)nb + 8;
        if (J->hres.bytes < need) {
            pin_put(J->hres);
            if ((r = pin_take(J->hres, need))) return r;
        }
    }
    // group maps
    std::vector<int32_t>& gm = J->hgm;
    gm.clear();
    int64_t ou_p = -1, ou_i = -1, ov_p = -1, ov_i = -1;
    if (d.groups_u_ptr) {
        const int64_t n = d.natoms_per_mol_u, m = d.groups_u_ptr[n];
        ou_p = gm.size(); gm.insert(gm.end(), d.groups_u_ptr, d.groups_u_ptr + n + 1);
        ou_i = gm.size(); gm.insert(gm.end(), d.groups_u_idx, d.groups_u_idx + m); gm.push_back(0);
    }
    if (d.groups_v_ptr) {
        const int64_t n = Nv, m = d.groups_v_ptr[n];
        ov_p = gm.size(); gm.insert(gm.end(), d.groups_v_ptr, d.groups_v_ptr + n + 1);
        ov_i = gm.size(); gm.insert(gm.end(), d.groups_v_idx, d.groups_v_idx + m); gm.push_back(0);
    }
    if (!gm.empty()) {
        if ((r = ensure(J, J->gmap, 4 * gm.size()))) return r;
        HIPCHK(hipMemcpyAsync(J->gmap.p, gm.data(), 4 * gm.size(), hipMemcpyHostToDevice, J->stream));
        HIPCHK(hipStreamSynchronize(J->stream));
        int32_t* base = (int32_t*)J->gmap.p;
        if (ou_p >= 0) { P.gu_ptr = base + ou_p; P.gu_idx = base + ou_i; }
        if (ov_p >= 0) { P.gv_ptr = base + ov_p; P.gv_idx = base + ov_i; }
    }
    J->d.groups_u_ptr = J->d.groups_u_idx = J->d.groups_v_ptr = J->d.groups_v_idx = nullptr;
    return CM_OK;
}

// The geometry error bits a job's batches left on the device (k_geo_flags):
// bit 0 molecules too large for the cell of the molecule-pair kernels, bit 1
// a cell grid larger than its allocation.  The same decoding for cm_job_finish
// and for the request server.
int geometry_error(uint64_t gerr)
{
    if (gerr & 1)
        return fail(CM_EUNSUPPORTED, "molecules too large for the unit cell: cutoff plus molecule radii exceed a plane spacing");
    if (gerr)
        return fail(CM_EHIP, "internal error: cell grid larger than its allocation");
    return CM_OK;
}

// Copies the read-back in h into the caller's arrays: FP64 accumulators, or
// 32-bit hits (packed) turned into (double)hits * w (w/2 in the self-MDDF
// group range), the product k_flush forms.  Worker threads take tasks of
// 256K entries; with ev, each first waits for the piece holding its task.
std::atomic<uint64_t>& pack_limit()
{
    static std::atomic<uint64_t> v{0xffffffffull};
    return v;
}

// Layout of a read-back in page-locked memory, for n accumulator entries of
// which the first nh = 4 nbins are the md/rdf histograms: the wire data
// (FP64 accumulators, or the 16-bit group counts of a packed read-back) in
// [0, 8 n), the cumulative md/rdf hits (64-bit), three flag words (geometry
// error bits, pack limit exceeded, number of overflowing counts), then the
// overflow list (entry, count) of a packed read-back.
struct RbLayout {
    int64_t n, nh, ng, ovcap;
    size_t off_hits, off_flags, off_ov, need, dev_ov;
    RbLayout(int64_t n_, int64_t nbins)
    {
        n = n_;
        nh = 4 * nbins;
        ng = n - nh;
        off_hits = 8 * (size_t)n;
        off_flags = off_hits + 8 * (size_t)nh;
        off_ov = off_flags + 32;
        // the overflow list sits in the accumulator block after the 16-bit words
        dev_ov = ((size_t)2 * ng + 15) & ~(size_t)15;
        ovcap = std::min<int64_t>(4096, (int64_t)((8 * (size_t)n - dev_ov) / 16));
        need = off_ov + 16 * (size_t)ovcap;
    }
};

// Writes a read-back (layout L in h) into the caller's arrays: FP64 copies,
// or for a packed read-back the products (double)count * w (w/2 in the
// self-MDDF group range [half_lo, half_hi)), the values k_flush forms; worker
// threads take tasks of 256K entries, and with ev each first waits for the
// wire piece holding its task.
int copy_out(const cm_job_desc& d, const RbLayout& L, const char* h, bool packed, double w, int64_t half_lo,
             int64_t half_hi, size_t piece, const hipEvent_t* ev, cm_raw* out)
{
    const int64_t nb = d.nbins, ngu = d.n_groups_u, ngv = d.n_groups_v;
    const int64_t n = L.n, nh = L.nh;
    const double wh = w / 2;
    auto mult = [&](int64_t k) { return (k >= half_lo && k < half_hi) ? wh : w; };
    // the caller's arrays as entry ranges of the accumulator block, in order
    struct Seg { int64_t o, n; double* dst; };
    std::vector<Seg> segs;
    int64_t o = 0;
    auto add = [&](double* dst, int64_t m) { if (dst && m) segs.push_back({o, m, dst}); o += m; };
    add(out->md_count, nb);
    add(out->md_count_random, nb);
    add(out->rdf_count, nb);
    add(out->rdf_count_random, nb);
    add(out->solute_group_count, ngu * nb);
    add(out->solvent_group_count, ngv * nb);
    add(out->solute_group_count_random, ngu * nb);
    add(out->solvent_group_count_random, ngv * nb);
    const ull* hits = (const ull*)(h + L.off_hits);
    const uint16_t* w16 = (const uint16_t*)h;
    constexpr int64_t TASK = (int64_t)1 << 18;
    const int64_t ntask = (n + TASK - 1) / TASK;
    hipError_t werr = hipSuccess;
    std::mutex werr_mu;
    auto work = [&](int64_t t0, int64_t step) {
        for (int64_t t = t0; t < ntask; t += step) {
            const int64_t a = t * TASK, b = std::min(n, a + TASK);
            if (ev) {
                // pieces land in stream order: the piece holding the task's last byte
                const size_t last = packed ? (b > nh ? 2 * (size_t)(b - 1 - nh) + 1 : 0) : 8 * (size_t)b - 1;
                if (!packed || b > nh) {
                    const hipError_t e = hipEventSynchronize(ev[last / piece]);
                    if (e != hipSuccess) {
                        std::lock_guard<std::mutex> lk(werr_mu);
                        werr = e;
                        return;
                    }
                }
            }
            for (const Seg& sg : segs) {
                const int64_t lo = std::max(a, sg.o), hi = std::min(b, sg.o + sg.n);
                if (lo >= hi) continue;
                double* dst = sg.dst + (lo - sg.o);
                if (!packed) {
                    memcpy(dst, h + 8 * lo, 8 * (size_t)(hi - lo));
                    continue;
                }
                // one multiplier per array (the self-MDDF half range is whole
                // arrays); signed conversions, a single instruction each
                const double m = mult(lo);
                if (hi <= nh) {
                    for (int64_t k = lo; k < hi; k++) dst[k - lo] = (double)(int64_t)hits[k] * m;
                } else {
                    const uint16_t* src = w16 + (lo - nh);
                    const int64_t cnt = hi - lo;
                    for (int64_t k = 0; k < cnt; k++) dst[k] = (double)(int32_t)src[k] * m;
                }
            }
        }
    };
    const int64_t nt = std::min<int64_t>(8, ntask);
    std::vector<std::thread> th;
    for (int64_t q = 1; q < nt; q++) th.emplace_back(work, q, nt);
    if (nt > 0) work(0, std::max<int64_t>(nt, 1));
    for (auto& x : th) x.join();
    if (werr != hipSuccess) return fail(CM_EHIP, std::string("read-back: ") + hipGetErrorString(werr));
    if (ev) HIPCHK(hipEventSynchronize(ev[RB_PIECES - 1]));
    if (packed) {
        // the counts that did not fit 16 bits
        uint64_t nov;
        memcpy(&nov, h + L.off_flags + 16, 8);
        const ulonglong2* ov = (const ulonglong2*)(h + L.off_ov);
        for (uint64_t q = 0; q < nov; q++) {
            const int64_t k = (int64_t)ov[q].x;
            for (const Seg& sg : segs)
                if (k >= sg.o && k < sg.o + sg.n) sg.dst[k - sg.o] = (double)ov[q].y * mult(k);
        }
    }
    auto cpu = [&](uint64_t* dst, const ull* src) { if (dst) memcpy(dst, src, 8 * nb); };
    cpu(out->hits_md, hits);
    cpu(out->hits_md_random, hits + nb);
    cpu(out->hits_rdf, hits + 2 * nb);
    cpu(out->hits_rdf_random, hits + 3 * nb);
    return CM_OK;
}

// The packed read-back of a one-weight job: k_pack16 into the accumulator
// block (its FP64 contents are not needed: the job has no flushed run)
void launch_pack16(cm_job* J, const RbLayout& L, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack16, dim3(grid_for(L.n)), dim3(BLOCK), 0, s, (const ull*)J->cnt.p, (uint16_t*)J->acc.p,
                       (ull*)J->hits.p, L.n, L.nh, (ull)pack_limit().load(), (ull*)J->stats.p + 3,
                       (ulonglong2*)((char*)J->acc.p + L.dev_ov), L.ovcap);
}

// a packed read-back that cannot be decoded: a count above the pack limit,
// or more overflowing counts than the list holds
bool pack_failed(const RbLayout& L, const char* h)
{
    uint64_t fl[3];
    memcpy(fl, h + L.off_flags, 24);
    return fl[1] != 0 || fl[2] > (uint64_t)L.ovcap;
}

// Read-back of cm_job_finish, pipelined: the accumulators come back in
// RB_PIECES pieces, each with an event, and worker threads copy 256K-entry
// tasks into the caller's arrays as soon as the piece holding a task has
// landed, so the copies into (often freshly allocated, page-faulting) caller
// memory overlap the transfer.  The geometry error bits come back first.
//
// When every frame of the job had one weight w (no weight change flushed
// yet), each FP64 result is (double)hits * w (w/2 in the self-MDDF group
// range) with nothing added before it, so the group counts cross PCIe as
// 16-bit words (a quarter of the bytes of the accumulators; the few larger
// counts in a list beside them, the md/rdf histograms from the 64-bit hits)
// and the host forms the same products, rounded the same way.  A count above
// the pack limit, or more large counts than the list holds, sends the job
// down the FP64 path.
int job_finish_pipelined(cm_job* J, cm_raw* out)
{
    const cm_job_desc& d = J->d;
    const int64_t nb = d.nbins;
    const RbLayout L(J->acc_len, nb);
    hipStream_t s = J->stream;
    const bool packed = J->pending && J->flushes == 0;
    const double w = J->pending_w;
    if (packed) {
        launch_pack16(J, L, s);
        HIPCHK(hipGetLastError());
    } else if (J->pending) {
        if (int r = flush(J, J->pending_w)) return r;
    }
    J->pending = false;
    if (J->hres.bytes < L.need) {
        pin_put(J->hres);
        int r = pin_take(J->hres, L.need);
        if (r) return r;
    }
    for (auto& e : J->rb_ev)
        if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    char* h = (char*)J->hres.p;
    // error bits and the packing flags, the md/rdf hits, the overflow list
    HIPCHK(hipMemcpyAsync(h + L.off_flags, (char*)J->stats.p + 16, 24, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(h + L.off_hits, J->hits.p, 8 * (size_t)L.nh, hipMemcpyDeviceToHost, s));
    if (packed)
        HIPCHK(hipMemcpyAsync(h + L.off_ov, (char*)J->acc.p + L.dev_ov, 16 * (size_t)L.ovcap, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(J->rb_ev[RB_PIECES], s));
    const size_t wire = packed ? 2 * (size_t)L.ng : 8 * (size_t)L.n;
    const size_t piece = ((wire + RB_PIECES - 1) / RB_PIECES + 4095) & ~(size_t)4095;
    for (int k = 0; k < RB_PIECES; k++) {
        const size_t o = std::min(wire, k * piece), m = std::min(wire, o + piece) - o;
        if (m) HIPCHK(hipMemcpyAsync(h + o, (char*)J->acc.p + o, m, hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(J->rb_ev[k], s));
    }
    HIPCHK(hipEventSynchronize(J->rb_ev[RB_PIECES]));
    uint64_t gerr;
    memcpy(&gerr, h + L.off_flags, 8);
    if (int r = geometry_error(gerr)) return r;
    int64_t half_lo, half_hi;
    half_range(J, &half_lo, &half_hi);
    int r;
    if (packed && pack_failed(L, h)) {
        // the FP64 accumulators after all (the hits were already added to
        // the cumulative md/rdf ones)
        HIPCHK(hipStreamSynchronize(s));
        if ((r = flush(J, w, nullptr, true, false))) return r;
        HIPCHK(hipMemcpyAsync(h, J->acc.p, 8 * (size_t)L.n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        r = copy_out(d, L, h, false, w, half_lo, half_hi, piece, nullptr, out);
    } else {
        r = copy_out(d, L, h, packed, w, half_lo, half_hi, piece, J->rb_ev, out);
    }
    if (r) return r;
    out->volume_total = J->volume_total;
    J->finished = true;
    return CM_OK;
}

// Reads a finished request's results back into J->hres (layout RbLayout):
// the accumulators, or, when its frames shared one weight, the packed counts
// (*packed set, as in cm_job_finish), the md/rdf hits and the flags.
int job_collect(cm_job* J, bool* packed)
{
    const RbLayout L(J->acc_len, J->d.nbins);
    *packed = J->pending && J->flushes == 0;
    if (*packed) {
        launch_pack16(J, L, J->stream);
        HIPCHK(hipGetLastError());
    } else if (J->pending) {
        if (int r = flush(J, J->pending_w)) return r;
    }
    J->pending = false;
    if (J->hres.bytes < L.need) {
        pin_put(J->hres);
        int r = pin_take(J->hres, L.need);
        if (r) return r;
    }
    char* h = (char*)J->hres.p;
    HIPCHK(hipMemcpyAsync(h, J->acc.p, *packed ? 2 * (size_t)L.ng : 8 * (size_t)L.n, hipMemcpyDeviceToHost, J->stream));
    HIPCHK(hipMemcpyAsync(h + L.off_hits, J->hits.p, 8 * (size_t)L.nh, hipMemcpyDeviceToHost, J->stream));
    HIPCHK(hipMemcpyAsync(h + L.off_flags, (char*)J->stats.p + 16, 24, hipMemcpyDeviceToHost, J->stream));
    if (*packed)
        HIPCHK(hipMemcpyAsync(h + L.off_ov, (char*)J->acc.p + L.dev_ov, 16 * (size_t)L.ovcap, hipMemcpyDeviceToHost,
                              J->stream));
    HIPCHK(hipStreamSynchronize(J->stream));
    uint64_t gerr;
    memcpy(&gerr, h + L.off_flags, 8);
    if (int r = geometry_error(gerr)) return r;
    if (*packed && pack_failed(L, h)) {
        // FP64 accumulators (md/rdf hits already added)
        if (int r = flush(J, J->pending_w, nullptr, true, false)) return r;
        HIPCHK(hipMemcpyAsync(h, J->acc.p, 8 * (size_t)L.n, hipMemcpyDeviceToHost, J->stream));
        HIPCHK(hipStreamSynchronize(J->stream));
        *packed = false;
    }
    J->finished = true;
    return CM_OK;
}

// raw accumulators of a collected request into the caller's cm_raw
int fill_raw(const cm_job_desc& d, int64_t acc_len, const PinBuf& res, double volume, bool packed, double w,
             cm_raw* out)
{
    const RbLayout L(acc_len, d.nbins);
    const int64_t half_lo = d.autocorrelation ? L.nh : 0, half_hi = d.autocorrelation ? acc_len : 0;
    if (int r = copy_out(d, L, (const char*)res.p, packed, w, half_lo, half_hi, 1, nullptr, out)) return r;
    out->volume_total = volume;
    return CM_OK;
}

// Times of the batches profiled since the last call: waits for the job's
// stream, reads the event pairs and adds them to the job's totals.
int prof_resolve(cm_job* J)
{
    if (J->prof_done == J->prof_used) return CM_OK;
    HIPCHK(hipStreamSynchronize(J->stream));
    for (size_t b = J->prof_done; b < J->prof_used; b++) {
        const auto& e = J->prof_ev[b];
        const bool rnd = J->prof_random[b] != 0;
        float real = 0.f, ran = -1.f, tot = 0.f;
        HIPCHK(hipEventElapsedTime(&real, e[0], e[1]));
        if (rnd) HIPCHK(hipEventElapsedTime(&ran, e[2], e[3]));
        HIPCHK(hipEventElapsedTime(&tot, e[4], rnd ? e[3] : e[1]));
        J->eval_ms += real + (rnd ? ran : 0.f);
        J->eval_launches += rnd ? 2 : 1;
        J->total_ms += tot;
        J->total_launches += 1;
        J->launch_ms.emplace_back(real, ran);
    }
    J->prof_done = J->prof_used;
    return CM_OK;
}

}  // namespace

extern "C" {

const char* cm_last_error(void) { return g_err.c_str(); }

int cm_job_create(const cm_job_desc* desc, cm_job** out)
{
    if (!desc || !out) return fail(CM_EINVAL, "null argument");
    *out = nullptr;
    int r = validate_desc(*desc);
    if (r) return r;
    cm_job* J = new cm_job();
    if ((r = stream_take(&J->stream))) {
        cm_job_destroy(J);
        return r;
    }
    if ((r = job_setup(J, *desc))) {
        cm_job_destroy(J);
        return r;
    }
    *out = J;
    return CM_OK;
}

int cm_job_submit(cm_job* job, int32_t nframes, const int64_t* frame_index, const void* x_solute, const void* x_solvent,
                  int32_t dtype, const double* cells, const double* weights)
{
    return submit_common(job, nframes, frame_index, x_solute, x_solvent, dtype, cells, weights, true);
}

int cm_job_submit_device(cm_job* job, int32_t nframes, const int64_t* frame_index, const void* d_x_solute,
                         const void* d_x_solvent, int32_t dtype, const double* cells, const double* weights)
{
    return submit_common(job, nframes, frame_index, d_x_solute, d_x_solvent, dtype, cells, weights, false);
}

int cm_job_finish(cm_job* J, cm_raw* out)
{
    if (!J || !out) return fail(CM_EINVAL, "null argument");
    if (J->finished) return fail(CM_ESTATE, "job already finished");
    return job_finish_pipelined(J, out);
}

int cm_job_set_profiling(cm_job* J, int32_t on)
{
    if (!J) return fail(CM_EINVAL, "null job");
    J->profiling = on != 0;
    return CM_OK;
}

int cm_job_set_tuning(cm_job* J, double cell_fraction, int32_t columns, int32_t eval_workgroups)
{
    if (!J) return fail(CM_EINVAL, "null job");
    if (J->frames > 0) return fail(CM_ESTATE, "tuning must be set before the first submit");
    if (!(cell_fraction >= 0.2 && cell_fraction <= 1.0)) return fail(CM_EINVAL, "cell_fraction must be in [0.2, 1]");
    if (columns != 0 && columns != 1 && columns != 2 && columns != 4 && columns != 8)
        return fail(CM_EINVAL, "columns must be 0 (default), 1, 2, 4 or 8");
    if (columns == 0) columns = J->tile_c;
    if (!J->kp.mol_lane && (64 / columns) / J->kp.nv < 1) return fail(CM_EINVAL, "too many columns for the solvent molecule size");
    if (eval_workgroups < 0) return fail(CM_EINVAL, "eval_workgroups must be >= 0");
    J->cell_frac = cell_fraction;
    J->tile_c = columns;
    J->kp.mpt = J->kp.mol_lane ? 64 / columns : (64 / columns) / J->kp.nv;
    J->eval_grid = eval_workgroups;
    return CM_OK;
}

int cm_job_kernel_stats(cm_job* J, cm_kernel_stats* out)
{
    if (!J || !out) return fail(CM_EINVAL, "null argument");
    if (int r = prof_resolve(J)) return r;
    ull st[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(st, J->stats.p, 16, hipMemcpyDeviceToHost, J->stream));
    HIPCHK(hipStreamSynchronize(J->stream));
    out->eval_ms = J->eval_ms;
    out->eval_launches = J->eval_launches;
    out->total_ms = J->total_ms;
    out->total_launches = J->total_launches;
    out->frames = J->frames;
    out->candidate_pairs = -1;
    out->fp64_rechecks = (int64_t)st[0];
    return CM_OK;
}

int cm_job_launch_times(cm_job* J, int32_t cap, double* real_ms, double* random_ms, int32_t* n)
{
    if (!J || !n || (cap > 0 && (!real_ms || !random_ms))) return fail(CM_EINVAL, "null argument");
    if (int r = prof_resolve(J)) return r;
    *n = (int32_t)J->launch_ms.size();
    for (int32_t k = 0; k < cap && k < *n; k++) {
        real_ms[k] = J->launch_ms[k].first;
        random_ms[k] = J->launch_ms[k].second;
    }
    return CM_OK;
}

int cm_job_set_verify(cm_job* J, int32_t on)
{
    if (!J) return fail(CM_EINVAL, "null job");
    if (J->frames > 0) return fail(CM_ESTATE, "verification must be set before the first submit");
    if (J->arena) return fail(CM_EUNSUPPORTED, "no verification on the no-allocation path");
    J->verify = on != 0;
    J->kp.place_all = J->verify ? 1 : 0;
    return CM_OK;
}

int cm_job_records(cm_job* J, int64_t frame, int32_t pass, cm_md_record* out)
{
    if (!J || !out) return fail(CM_EINVAL, "null argument");
    if (!J->verify) return fail(CM_ESTATE, "verification is off for this job");
    if (!J->atoms_family)
        return fail(CM_EUNSUPPORTED, "records are kept for jobs with one solute molecule (tile kernels) only");
    if (pass < 0 || pass > J->K) return fail(CM_EINVAL, "pass must be 0 (real) or 1..n_random_samples");
    auto it = J->vstore.find(frame);
    if (it == J->vstore.end()) return fail(CM_EINVAL, "frame not processed by this job");
    static_assert(sizeof(VRec) == sizeof(cm_md_record), "record layout");
    std::memcpy(out, it->second.rec.data() + (size_t)pass * J->d.nmols_v, sizeof(VRec) * (size_t)J->d.nmols_v);
    return CM_OK;
}

int cm_job_placement(cm_job* J, int64_t frame, int32_t pass, double* out, uint8_t* placed)
{
    if (!J || !out || !placed) return fail(CM_EINVAL, "null argument");
    if (!J->verify) return fail(CM_ESTATE, "verification is off for this job");
    if (pass < 1 || pass > J->K) return fail(CM_EINVAL, "pass must be 1..n_random_samples");
    auto it = J->vstore.find(frame);
    if (it == J->vstore.end()) return fail(CM_EINVAL, "frame not processed by this job");
    const int64_t k = pass - 1;
    std::memcpy(out, it->second.placed.data() + (size_t)k * J->Nv * 3, sizeof(double) * 3 * (size_t)J->Nv);
    std::memcpy(placed, it->second.placed_flag.data() + (size_t)k * J->d.nmols_v, (size_t)J->d.nmols_v);
    return CM_OK;
}

int cm_job_reference_draws(cm_job* J, int64_t frame, int32_t* out)
{
    if (!J || !out) return fail(CM_EINVAL, "null argument");
    if (!J->verify) return fail(CM_ESTATE, "verification is off for this job");
    auto it = J->vstore.find(frame);
    if (it == J->vstore.end()) return fail(CM_EINVAL, "frame not processed by this job");
    std::memcpy(out, it->second.draws.data(), sizeof(int32_t) * (size_t)J->K);
    return CM_OK;
}

void cm_job_destroy(cm_job* J)
{
    if (!J) return;
    if (J->stream) (void)hipStreamSynchronize(J->stream);
    DevBuf* bufs[] = {&J->cnt, &J->acc, &J->hits, &J->stats, &J->gmap, &J->geo, &J->cells_cnt, &J->cells_start,
                      &J->tmp4, &J->tmpcell, &J->sorted, &J->solw, &J->ucen, &J->vw, &J->vcell, &J->vcen, &J->P,
                      &J->vw_r, &J->vcen_r, &J->bflag, &J->blist, &J->nbulk, &J->uk, &J->drawn_ptr,
                      &J->drawn_idx, &J->xin_u, &J->xin_v, &J->near, &J->mkey, &J->mrank, &J->mcnt, &J->mstart,
                      &J->msorted, &J->nnear, &J->chunk_base, &J->vrec};
    for (DevBuf* b : bufs) dev_put(*b);
    pin_put(J->hres);
    for (auto& a : J->prof_ev)
        for (auto& e : a) (void)hipEventDestroy(e);
    for (auto& e : J->rb_ev)
        if (e) (void)hipEventDestroy(e);
    stream_put(J->stream);
    delete J;
}

int cm_set_readback_pack_limit(uint64_t limit)
{
    if (limit > 0xffffffffull) return fail(CM_EINVAL, "the pack limit is at most 2^32 - 1");
    pack_limit().store(limit);
    return CM_OK;
}

int cm_release_cached_memory(void)
{
    for (auto& blk : cache_drain(dev_cache())) HIPCHK(hipFree(blk.second));
    for (auto& blk : cache_drain(pin_cache())) HIPCHK(hipHostFree(blk.second));
    return CM_OK;
}

int cm_device_alloc(size_t bytes, void** dptr)
{
    if (!dptr) return fail(CM_EINVAL, "null pointer");
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(CM_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return CM_OK;
}

int cm_device_upload(void* dptr, const void* host, size_t bytes)
{
    HIPCHK(hipMemcpy(dptr, host, bytes, hipMemcpyHostToDevice));
    return CM_OK;
}

int cm_device_free(void* dptr)
{
    HIPCHK(hipFree(dptr));
    return CM_OK;
}

int cm_device_synchronize(void)
{
    HIPCHK(hipDeviceSynchronize());
    return CM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// request server: slots (one stream and one pooled job each) take requests in
// arrival order when the footprint estimates of the admitted ones fit the budget
// ---------------------------------------------------------------------------
namespace {

struct SrvReq {
    uint64_t id = 0;
    cm_job_desc d{};
    std::vector<int32_t> gup, gui, gvp, gvi;
    std::vector<int64_t> fi;
    std::vector<double> cells, w;
    const void* xu = nullptr;
    const void* xv = nullptr;
    int32_t nframes = 0, dtype = 0;
    double cf = 0.0;
    int32_t cols = 0;
    int64_t foot = 0;
    bool done = false;
    bool waited = false;             // a cm_server_wait has taken the ticket
    int status = CM_OK;
    std::string err;
    PinBuf res;                     // read-back accumulators (taken over from the slot's job)
    int64_t acc_len = 0;
    bool packed = false;            // res holds 32-bit hits of frames of weight wt
    double wt = 0.0;
    double volume = 0.0;
    cm_request_times t{};
    ~SrvReq() { pin_put(res); }
};

int64_t job_footprint(const cm_job_desc& d, int32_t nframes, int32_t dtype, const double* cells, int64_t* bytes)
{
    cm_job J;
    job_params(&J, d);
    int64_t cap, kcap, Fmax, pfb;
    int r = batch_plan(&J, nframes, cells, dtype == CM_F32 ? 4 : 8, false, &cap, &kcap, &Fmax, &pfb);
    if (r) return r;
    int64_t groups = 0;
    if (d.groups_u_ptr) groups += 4 * ((int64_t)d.natoms_per_mol_u + 2 + d.groups_u_ptr[d.natoms_per_mol_u]);
    if (d.groups_v_ptr) groups += 4 * (J.Nv + 2 + d.groups_v_ptr[J.Nv]);
    *bytes = 16 * J.acc_len + 32 * (int64_t)d.nbins + 64 + groups + std::min<int64_t>(Fmax, nframes) * pfb;
    return CM_OK;
}

}  // namespace

struct cm_server {
    int slots = 4;
    int64_t budget = 0;
    int device = 0;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<SrvReq*> queue;
    std::unordered_map<uint64_t, SrvReq*> reqs;
    int64_t reserved = 0;
    bool stop = false;
    uint64_t next = 1;
    std::vector<cm_job*> shells;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

int run_request(cm_job* J, SrvReq* q)
{
    cm_job_desc d = q->d;
    d.groups_u_ptr = q->gup.empty() ? nullptr : q->gup.data();
    d.groups_u_idx = q->gup.empty() ? nullptr : q->gui.data();
    d.groups_v_ptr = q->gvp.empty() ? nullptr : q->gvp.data();
    d.groups_v_idx = q->gvp.empty() ? nullptr : q->gvi.data();
    int r = job_setup(J, d);
    if (r) return r;
    if (q->cf > 0.0 || q->cols > 0)
        if ((r = cm_job_set_tuning(J, q->cf > 0.0 ? q->cf : 0.5, q->cols, 0))) return r;
    if ((r = submit_common(J, q->nframes, q->fi.data(), q->xu, q->xv, q->dtype, q->cells.data(), q->w.data(), false)))
        return r;
    const double w = J->pending_w;
    bool packed = false;
    if ((r = job_collect(J, &packed))) return r;
    q->packed = packed;
    q->wt = w;
    q->res = J->hres;
    J->hres = PinBuf{};
    q->acc_len = J->acc_len;
    q->volume = J->volume_total;
    return CM_OK;
}

void serve_slot(cm_server* S, int k)
{
    // the device current when the server was created (hipGetDevice returned it)
    (void)hipSetDevice(S->device);
    cm_job* J = S->shells[k];
    for (;;) {
        SrvReq* q;
        {
            std::unique_lock<std::mutex> lk(S->m);
            S->cv_work.wait(lk, [&] {
                if (S->queue.empty()) return S->stop;
                // arrival order; a request larger than the budget runs alone
                return S->reserved == 0 || S->reserved + S->queue.front()->foot <= S->budget;
            });
            if (S->queue.empty()) return;
            q = S->queue.front();
            S->queue.pop_front();
            S->reserved += q->foot;
            q->t.started = S->now();
        }
        const int r = run_request(J, q);
        {
            std::lock_guard<std::mutex> lk(S->m);
            S->reserved -= q->foot;
            q->status = r;
            if (r) q->err = g_err;
            q->t.finished = S->now();
            q->done = true;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

int cm_job_footprint(const cm_job_desc* desc, int32_t nframes, int32_t dtype, const double* cells, int64_t* bytes)
{
    if (!desc || !bytes || (nframes > 0 && !cells)) return fail(CM_EINVAL, "null argument");
    if (nframes < 0) return fail(CM_EINVAL, "nframes must be >= 0");
    if (dtype != CM_F64 && dtype != CM_F32) return fail(CM_EINVAL, "dtype must be CM_F64 or CM_F32");
    int r = validate_desc(*desc);
    if (r) return r;
    return (int)job_footprint(*desc, nframes, dtype, cells, bytes);
}

int cm_server_create(const cm_server_desc* desc, cm_server** out)
{
    if (!out) return fail(CM_EINVAL, "null argument");
    *out = nullptr;
    const int slots = desc && desc->slots > 0 ? desc->slots : 4;
    if (slots > 64) return fail(CM_EINVAL, "at most 64 slots");
    cm_server* S = new cm_server();
    S->slots = slots;
    S->budget = desc && desc->memory_budget > 0 ? desc->memory_budget : (int64_t)200 << 30;
    S->t0 = std::chrono::steady_clock::now();
    if (hipGetDevice(&S->device) != hipSuccess) {
        delete S;
        return fail(CM_EHIP, "hipGetDevice failed");
    }
    for (int k = 0; k < slots; k++) {
        cm_job* J = new cm_job();
        if (hipStreamCreateWithFlags(&J->stream, hipStreamNonBlocking) != hipSuccess) {
            cm_job_destroy(J);
            cm_server_destroy(S);
            return fail(CM_EHIP, "hipStreamCreate failed");
        }
        S->shells.push_back(J);
    }
    for (int k = 0; k < slots; k++) S->workers.emplace_back(serve_slot, S, k);
    *out = S;
    return CM_OK;
}

int cm_server_enqueue(cm_server* S, const cm_request* req, uint64_t* ticket)
{
    if (!S || !req || !req->job || !ticket) return fail(CM_EINVAL, "null argument");
    const cm_job_desc& d = *req->job;
    int r = validate_desc(d);
    if (r) return r;
    if (req->nframes < 1 || !req->frame_index || !req->d_x_solute || !req->d_x_solvent || !req->cells || !req->weights)
        return fail(CM_EINVAL, "invalid frame arrays");
    if (req->dtype != CM_F64 && req->dtype != CM_F32) return fail(CM_EINVAL, "dtype must be CM_F64 or CM_F32");
    for (int f = 0; f < req->nframes; f++) {
        if (!(req->weights[f] > 0.0) || !std::isfinite(req->weights[f])) return fail(CM_EINVAL, "frame weights must be positive and finite");
        if (req->frame_index[f] < 1 || req->frame_index[f] > 0xFFFFFFFFll) return fail(CM_EINVAL, "frame_index must be in [1, 2^32)");
    }
    if (req->cell_fraction != 0.0 && !(req->cell_fraction >= 0.2 && req->cell_fraction <= 1.0))
        return fail(CM_EINVAL, "cell_fraction must be in [0.2, 1]");
    if (req->columns != 0 && req->columns != 1 && req->columns != 2 && req->columns != 4 && req->columns != 8)
        return fail(CM_EINVAL, "columns must be 0 (default), 1, 2, 4 or 8");
    SrvReq* q = new SrvReq();
    q->d = d;
    const int64_t Nv = (int64_t)d.nmols_v * d.natoms_per_mol_v;
    if (d.groups_u_ptr) {
        q->gup.assign(d.groups_u_ptr, d.groups_u_ptr + d.natoms_per_mol_u + 1);
        q->gui.assign(d.groups_u_idx, d.groups_u_idx + d.groups_u_ptr[d.natoms_per_mol_u]);
    }
    if (d.groups_v_ptr) {
        q->gvp.assign(d.groups_v_ptr, d.groups_v_ptr + Nv + 1);
        q->gvi.assign(d.groups_v_idx, d.groups_v_idx + d.groups_v_ptr[Nv]);
    }
    q->d.groups_u_ptr = q->d.groups_u_idx = q->d.groups_v_ptr = q->d.groups_v_idx = nullptr;
    q->fi.assign(req->frame_index, req->frame_index + req->nframes);
    q->cells.assign(req->cells, req->cells + 9 * (size_t)req->nframes);
    q->w.assign(req->weights, req->weights + req->nframes);
    q->xu = req->d_x_solute;
    q->xv = req->d_x_solvent;
    q->nframes = req->nframes;
    q->dtype = req->dtype;
    q->cf = req->cell_fraction;
    q->cols = req->columns;
    if ((r = (int)job_footprint(d, req->nframes, req->dtype, req->cells, &q->foot))) {
        delete q;
        return r;
    }
    {
        std::lock_guard<std::mutex> lk(S->m);
        if (S->stop) {
            delete q;
            return fail(CM_ESTATE, "server is shutting down");
        }
        q->id = S->next++;
        q->t.enqueued = S->now();
        S->reqs[q->id] = q;
        S->queue.push_back(q);
        *ticket = q->id;
    }
    S->cv_work.notify_all();
    return CM_OK;
}

int cm_server_wait(cm_server* S, uint64_t ticket, cm_raw* out, cm_request_times* times)
{
    if (!S) return fail(CM_EINVAL, "null server");
    SrvReq* q;
    {
        std::unique_lock<std::mutex> lk(S->m);
        auto it = S->reqs.find(ticket);
        if (it == S->reqs.end() || it->second->waited) return fail(CM_EINVAL, "unknown ticket");
        q = it->second;
        // a second waiter on the same ticket is refused while this one waits
        q->waited = true;
        // the map may rehash while the lock is released inside wait (an
        // enqueue on another thread inserts): the entry is erased by key
        S->cv_done.wait(lk, [&] { return q->done; });
        S->reqs.erase(ticket);
    }
    if (times) *times = q->t;
    int st = q->status;
    if (st == CM_OK) {
        if (out) st = fill_raw(q->d, q->acc_len, q->res, q->volume, q->packed, q->wt, out);
    } else {
        g_err = q->err;
    }
    delete q;
    return st;
}

void cm_server_destroy(cm_server* S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    for (cm_job* J : S->shells) cm_job_destroy(J);
    for (auto& kv : S->reqs) delete kv.second;
    delete S;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// many jobs per round (cm_jobs_submit_device)
// ---------------------------------------------------------------------------
namespace {

constexpr int MR_MAX_DEVICES = 64;
constexpr int MS_FMAX = 48;        // frames of one job per round (the single-job batch cap)
// Blocks of the stage launches.  The slot stages have one virtual block per
// (256 slots, frame, pass) of every job, most of which exit at once (their
// slots are not near); a block walks its virtual blocks one after another,
// so the grids are large enough that each walks a few dozen at most (a grid
// of 2048 took 1.8x the single-job launches' time in the near placement).
constexpr int MG_256 = 32768;      // blocks of the 256-thread stage launches
constexpr int MG_1024 = 4096;      // blocks of the 1024-thread stage launches
constexpr int MG_64 = 1024;        // blocks of the 64-thread stage launches
constexpr int64_t MULTI_BUDGET = (int64_t)64 << 30;   // batch buffers of one submission's jobs together

struct MultiRunner {
    std::mutex m;
    bool ready = false;
    hipStream_t stream = nullptr;
    void* tab = nullptr;                  // device table and geometry area (fixed address: graphs use it)
    size_t tab_bytes = 0;
    void* htab[2] = {nullptr, nullptr};   // page-locked staging of the table, used in turn
    hipEvent_t up[2] = {nullptr, nullptr};
    int cur = 0;
    bool graphs = true;
    std::map<int, hipGraphExec_t> exec;   // per kernel variant
    hipEvent_t done = nullptr;
    int cus = 256;
    int64_t rounds = 0, launches = 0;
};

MultiRunner* multi_runners()
{
    static MultiRunner* r = new MultiRunner[MR_MAX_DEVICES];
    return r;
}

// upload layout: the table, the two evaluation launches' chunk counters
// (zero in every upload), the frames' Geo records
constexpr size_t MT_CTR_OFF = (sizeof(MultiTab) + 255) & ~(size_t)255;
constexpr size_t MT_GEO_OFF = MT_CTR_OFF + 256;

int multi_init(MultiRunner& R)
{
    if (R.ready) return CM_OK;
    HIPCHK(hipStreamCreateWithFlags(&R.stream, hipStreamNonBlocking));
    R.tab_bytes = MT_GEO_OFF + sizeof(Geo) * (size_t)MS_MAX * MS_FMAX;
    HIPCHK(hipMalloc(&R.tab, R.tab_bytes));
    for (int k = 0; k < 2; k++) {
        HIPCHK(hipHostMalloc(&R.htab[k], R.tab_bytes, hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&R.up[k], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&R.done, hipEventDisableTiming));
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess)
        if (hipDeviceGetAttribute(&R.cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || R.cus < 1) R.cus = 256;
    R.ready = true;
    return CM_OK;
}

template <class KERN>
int persistent_grid(KERN kern, size_t shm, int cus)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 64, shm) != hipSuccess || nb < 1) nb = 1;
    return nb * cus;
}

// the stage launches of one round (every argument but the table address is
// fixed per kernel variant, so the sequence can be captured and replayed)
template <class T, int C, int NVM>
void multi_launch(hipStream_t s, const MultiTab* tab, int tcap, int cus)
{
    const size_t shm = (sizeof(float4) + (NVM > 0 ? sizeof(float) : 0)) * (size_t)tcap;
    auto ev_real = km_eval_tiles<T, T, false, C, NVM>;
    auto ev_rand = km_eval_tiles<T, double, true, C, NVM>;
    // persistent grids of the two evaluations (occupancy queried once per
    // variant, before any capture)
    static const int g_real = persistent_grid(ev_real, shm, cus);
    static const int g_rand = persistent_grid(ev_rand, shm, cus);
    hipLaunchKernelGGL(km_zero, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_radii<T>, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_grid, dim3(MG_64), dim3(64), 0, s, tab);
    hipLaunchKernelGGL(km_wrap_atoms<T>, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_scan_cells, dim3(MG_1024), dim3(1024), 0, s, tab);
    hipLaunchKernelGGL(km_scatter, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_near_mark, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_dfield, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_prep_solvent_tiles<T>, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_scan_rows<false>, dim3(MG_1024), dim3(1024), 0, s, tab);
    hipLaunchKernelGGL(km_mol_scatter<false>, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_fill_real<T>, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_chunk_table<false>, dim3(MG_1024), dim3(1024), 0, s, tab);
    hipLaunchKernelGGL(km_bulk_init, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    int* ctr = (int*)((char*)tab + MT_CTR_OFF);
    hipLaunchKernelGGL(ev_real, dim3(g_real), dim3(64), shm, s, tab, ctr);
    hipLaunchKernelGGL(km_bulk_compact, dim3(MG_1024), dim3(1024), 0, s, tab);
    hipLaunchKernelGGL(km_place_keys, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_scan_rows<true>, dim3(MG_1024), dim3(1024), 0, s, tab);
    hipLaunchKernelGGL(km_mol_scatter<true>, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_place_near<T>, dim3(MG_256), dim3(BLOCK), 0, s, tab);
    hipLaunchKernelGGL(km_chunk_table<true>, dim3(MG_1024), dim3(1024), 0, s, tab);
    hipLaunchKernelGGL(ev_rand, dim3(g_rand), dim3(64), shm, s, tab, ctr + 1);
    hipLaunchKernelGGL(km_geo_flags, dim3(MG_64), dim3(64), 0, s, tab);
}

template <class T, int NVM>
void multi_launch_c(int c, hipStream_t s, const MultiTab* tab, int tcap, int cus)
{
    if (c == 8) multi_launch<T, 8, NVM>(s, tab, tcap, cus);
    else if (c == 4) multi_launch<T, 4, NVM>(s, tab, tcap, cus);
    else if (c == 2) multi_launch<T, 2, NVM>(s, tab, tcap, cus);
    else multi_launch<T, 1, NVM>(s, tab, tcap, cus);
}

void multi_launch_any(bool f32, int mol_lane, int c, hipStream_t s, const MultiTab* tab, int tcap, int cus)
{
    if (f32) {
        if (mol_lane) multi_launch_c<float, 3>(c, s, tab, tcap, cus);
        else multi_launch_c<float, 0>(c, s, tab, tcap, cus);
    } else {
        if (mol_lane) multi_launch_c<double, 3>(c, s, tab, tcap, cus);
        else multi_launch_c<double, 0>(c, s, tab, tcap, cus);
    }
}

// one job's part of a multi-job submission
struct MJob {
    cm_job* J = nullptr;
    const cm_frames* fr = nullptr;
    int64_t cap = 0, kcap = 0;
    std::vector<std::pair<int, int>> batches;   // (first frame, frames)
};

// blocks of each stage's grid for one job's batch of F frames, as the
// single-job launches size them: (x extent, total)
void stage_blocks(const KP& P, int F, const int64_t z16[3], int nbx[MS_N], int cnt[MS_N])
{
    auto set = [&](int st, int64_t x, int64_t y) {
        x = std::max<int64_t>(x, 1);
        nbx[st] = (int)x;
        cnt[st] = (int)(x * std::max<int64_t>(y, 0));
    };
    const int64_t mvb = (P.Mv + BLOCK - 1) / BLOCK;
    const int K = P.K;
    set(MS_ZERO, std::min<int64_t>(1024, std::max<int64_t>(1, (z16[0] + z16[1] + z16[2] + BLOCK - 1) / BLOCK)), 1);
    set(MS_RADII, std::min<int64_t>(mvb, 64), F);
    set(MS_GRID, (F + 63) / 64, 1);
    set(MS_WRAP, grid_for((int64_t)F * P.nu), 1);
    set(MS_SCAN_CELLS, F, 1);
    set(MS_SCATTER, grid_for((int64_t)F * P.nu), 1);
    set(MS_NEAR, (P.nu + BLOCK - 1) / BLOCK, std::min(F, 65535));
    set(MS_DFIELD, grid_for((int64_t)F * P.cap * 16), 1);
    set(MS_PREP, grid_for((int64_t)F * P.Mv), 1);
    set(MS_SCAN_ROWS0, F, 1);
    set(MS_MOL_SCATTER0, mvb, std::min(F, 65535));
    set(MS_FILL, grid_for((int64_t)F * P.Mv), 1);
    set(MS_CHUNK0, 1, 1);
    set(MS_BULK_INIT, grid_for((int64_t)F * P.kb * P.Mv), K > 0 ? 1 : 0);
    set(MS_BULK_COMPACT, (int64_t)F * P.kb, K > 0 ? 1 : 0);
    set(MS_PLACE_KEYS, mvb, K > 0 ? std::min(F * K, 65535) : 0);
    set(MS_SCAN_ROWSK, (int64_t)F * K, K > 0 ? 1 : 0);
    set(MS_MOL_SCATTERK, mvb, K > 0 ? std::min(F * K, 65535) : 0);
    set(MS_PLACE_NEAR, mvb, K > 0 ? std::min(F * K, 65535) : 0);
    set(MS_CHUNKK, 1, K > 0 ? 1 : 0);
    set(MS_GEO_FLAGS, 1, 1);
}

// the rounds of one group of jobs that share the kernel variant
int multi_run_group(MultiRunner& R, std::vector<MJob>& grp, int dtype)
{
    const int ds = dtype == CM_F32 ? 4 : 8;
    const cm_job* J0 = grp[0].J;
    const int variant = (dtype == CM_F32 ? 1 : 0) | (J0->kp.mol_lane ? 2 : 0) | (J0->tile_c << 2);
    size_t maxr = 0;
    for (auto& m : grp) maxr = std::max(maxr, m.batches.size());
    for (size_t r = 0; r < maxr; r++) {
        std::vector<MJob*> act;
        for (auto& m : grp)
            if (r < m.batches.size()) act.push_back(&m);
        for (size_t a0 = 0; a0 < act.size(); a0 += MS_MAX) {
            const int n = (int)std::min<size_t>(MS_MAX, act.size() - a0);
            // the staging buffer's previous upload must have been read
            HIPCHK(hipEventSynchronize(R.up[R.cur]));
            MultiTab* T = (MultiTab*)R.htab[R.cur];
            Geo* gh = (Geo*)((char*)R.htab[R.cur] + MT_GEO_OFF);
            Geo* gd = (Geo*)((char*)R.tab + MT_GEO_OFF);
            T->nsys = n;
            T->pad = 0;
            memset((char*)R.htab[R.cur] + MT_CTR_OFF, 0, MT_GEO_OFF - MT_CTR_OFF);
            int64_t goff = 0;
            int cnt[MS_N];
            for (int st = 0; st < MS_N; st++) T->base[st][0] = 0;
            for (int i = 0; i < n; i++) {
                MJob& m = *act[a0 + i];
                cm_job* J = m.J;
                const cm_frames& fr = *m.fr;
                const int f0 = m.batches[r].first, F = m.batches[r].second;
                const double w = fr.weights[f0];
                if (J->pending && J->pending_w != w)
                    if (int e = flush(J, J->pending_w, R.stream)) return e;
                J->pending = true;
                J->pending_w = w;
                for (int f = 0; f < F; f++) {
                    cm::Cell64 c;
                    cm::cell_setup(fr.cells + 9 * (f0 + f), J->d.triclinic, &c);
                    J->volume_total += w * cm::cell_volume(c.H);
                }
                if (int e = batch_geo(J, F, fr.frame_index + f0, fr.cells + 9 * f0, m.cap, m.kcap, gh + goff)) return e;
                SysDev& S = T->sys[i];
                memset(&S, 0, sizeof(SysDev));
                S.P = J->kp;
                S.P.F = F;
                S.P.cap = (int)m.cap;
                S.P.kcap = (int)m.kcap;
                S.B = batch_bufs(J, gd + goff);
                S.B.dfield = (unsigned*)((char*)J->near.p + near_flags_bytes(F, m.cap));
                S.xu = (const char*)fr.d_x_solute + (size_t)f0 * J->Nu * 3 * ds;
                S.xv = (const char*)fr.d_x_solvent + (size_t)f0 * J->Nv * 3 * ds;
                S.zp[0] = (uint4*)J->cells_cnt.p;
                S.zp[1] = (uint4*)J->near.p;
                S.zp[2] = (uint4*)J->mcnt.p;
                S.z16[0] = (int64_t)(round16(sizeof(int) * (size_t)F * (m.cap + 1)) / 16);
                S.z16[1] = (int64_t)(near_bytes(F, m.cap) / 16);
                S.z16[2] = (int64_t)(round16((size_t)4 * F * S.P.NP * (m.kcap + 1)) / 16);
                stage_blocks(S.P, F, S.z16, S.nbx, cnt);
                for (int st = 0; st < MS_N; st++) T->base[st][i + 1] = T->base[st][i] + cnt[st];
                goff += F;
                J->frames += F;
            }
            HIPCHK(hipMemcpyAsync(R.tab, R.htab[R.cur], MT_GEO_OFF + sizeof(Geo) * (size_t)goff, hipMemcpyHostToDevice,
                                  R.stream));
            HIPCHK(hipEventRecord(R.up[R.cur], R.stream));
            R.cur ^= 1;
            const bool f32 = dtype == CM_F32;
            if (R.graphs) {
                auto it = R.exec.find(variant);
                if (it == R.exec.end()) {
                    // first round of this variant: launched directly (which also
                    // settles the occupancy queries), captured from the next one
                    multi_launch_any(f32, J0->kp.mol_lane, J0->tile_c, R.stream, (const MultiTab*)R.tab, J0->kp.tcap,
                                     R.cus);
                    HIPCHK(hipGetLastError());
                    R.exec.emplace(variant, nullptr);
                    R.rounds++;
                    R.launches += MS_N + 2;
                    continue;
                }
                if (!it->second) {
                    hipGraph_t g = nullptr;
                    hipGraphExec_t x = nullptr;
                    HIPCHK(hipStreamBeginCapture(R.stream, hipStreamCaptureModeThreadLocal));
                    multi_launch_any(f32, J0->kp.mol_lane, J0->tile_c, R.stream, (const MultiTab*)R.tab, J0->kp.tcap, R.cus);
                    HIPCHK(hipStreamEndCapture(R.stream, &g));
                    HIPCHK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
                    HIPCHK(hipGraphDestroy(g));
                    it->second = x;
                }
                HIPCHK(hipGraphLaunch(it->second, R.stream));
            } else {
                multi_launch_any(f32, J0->kp.mol_lane, J0->tile_c, R.stream, (const MultiTab*)R.tab, J0->kp.tcap, R.cus);
                HIPCHK(hipGetLastError());
            }
            R.rounds++;
            R.launches += MS_N + 2;
        }
    }
    return CM_OK;
}

}  // namespace

extern "C" {

int cm_jobs_submit_device(int32_t njobs, cm_job* const* jobs, const cm_frames* frames)
{
    if (njobs < 0 || (njobs > 0 && (!jobs || !frames))) return fail(CM_EINVAL, "null argument");
    for (int k = 0; k < njobs; k++) {
        cm_job* J = jobs[k];
        const cm_frames& fr = frames[k];
        if (!J) return fail(CM_EINVAL, "null job");
        if (J->finished) return fail(CM_ESTATE, "job already finished");
        for (int q = 0; q < k; q++)
            if (jobs[q] == J) return fail(CM_EINVAL, "a job appears twice");
        if (fr.nframes < 0 || (fr.nframes > 0 && (!fr.frame_index || !fr.d_x_solute || !fr.d_x_solvent || !fr.cells ||
                                                  !fr.weights)))
            return fail(CM_EINVAL, "invalid frame arrays");
        if (fr.dtype != CM_F64 && fr.dtype != CM_F32) return fail(CM_EINVAL, "dtype must be CM_F64 or CM_F32");
        for (int f = 0; f < fr.nframes; f++) {
            if (!(fr.weights[f] > 0.0) || !std::isfinite(fr.weights[f])) return fail(CM_EINVAL, "frame weights must be positive and finite");
            if (fr.frame_index[f] < 1 || fr.frame_index[f] > 0xFFFFFFFFll) return fail(CM_EINVAL, "frame_index must be in [1, 2^32)");
        }
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MR_MAX_DEVICES) return fail(CM_EUNSUPPORTED, "device index beyond the supported range");
    MultiRunner& R = multi_runners()[dev];
    std::lock_guard<std::mutex> lk(R.m);
    if (int r = multi_init(R)) return r;
    // the previous submission's rounds may still use buffers that are about
    // to grow
    HIPCHK(hipStreamSynchronize(R.stream));
    // groups of jobs sharing the kernel variant; the others run on their own.
    // The jobs of a submission hold their batch buffers at once: each gets an
    // equal share of MULTI_BUDGET for its frames per round (at least one).
    int nmulti = 0;
    for (int k = 0; k < njobs; k++)
        if (frames[k].nframes > 0 && jobs[k]->atoms_family && !jobs[k]->verify && !jobs[k]->arena && !jobs[k]->profiling)
            nmulti++;
    const int64_t share = MULTI_BUDGET / std::max(1, nmulti);
    std::map<int, std::vector<MJob>> groups;
    std::map<int, int> group_dtype;
    for (int k = 0; k < njobs; k++) {
        cm_job* J = jobs[k];
        const cm_frames& fr = frames[k];
        if (fr.nframes == 0) continue;
        if (!J->atoms_family || J->verify || J->arena || J->profiling) {
            if (int r = submit_common(J, fr.nframes, fr.frame_index, fr.d_x_solute, fr.d_x_solvent, fr.dtype, fr.cells,
                                      fr.weights, false))
                return r;
            continue;
        }
        const int ds = fr.dtype == CM_F32 ? 4 : 8;
        MJob m;
        m.J = J;
        m.fr = &fr;
        int64_t Fmax, pfb;
        if (int r = batch_plan(J, fr.nframes, fr.cells, ds, false, &m.cap, &m.kcap, &Fmax, &pfb)) return r;
        Fmax = std::max<int64_t>(1, std::min<int64_t>({Fmax, (int64_t)MS_FMAX, share / std::max<int64_t>(pfb, 1)}));
        // batches as the single-job path cuts them: equal batches over each
        // run of one frame weight
        int Fbig = 1;
        for (int f0 = 0; f0 < fr.nframes;) {
            const double w = fr.weights[f0];
            int run = 1;
            while (f0 + run < fr.nframes && fr.weights[f0 + run] == w) run++;
            const int nb = (int)((run + Fmax - 1) / Fmax);
            const int F = (run + nb - 1) / nb;
            for (int b = 0; b < run; b += F) m.batches.emplace_back(f0 + b, std::min(F, run - b));
            Fbig = std::max(Fbig, F);
            f0 += run;
        }
        // the job's own stream may still run earlier work on these buffers
        HIPCHK(hipStreamSynchronize(J->stream));
        if (int r = alloc_batch(J, Fbig, m.cap, m.kcap, ds, false)) return r;
        const int key = (fr.dtype == CM_F32 ? 1 : 0) | (J->kp.mol_lane ? 2 : 0) | (J->tile_c << 2);
        groups[key].push_back(m);
        group_dtype[key] = fr.dtype;
    }
    for (auto& kv : groups)
        if (int r = multi_run_group(R, kv.second, group_dtype[kv.first])) return r;
    // the jobs' own streams (cm_job_finish, later submits) wait for the rounds
    HIPCHK(hipEventRecord(R.done, R.stream));
    for (auto& kv : groups)
        for (auto& m : kv.second) HIPCHK(hipStreamWaitEvent(m.J->stream, R.done, 0));
    return CM_OK;
}

int cm_multi_set_graphs(int32_t on)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MR_MAX_DEVICES) return fail(CM_EUNSUPPORTED, "device index beyond the supported range");
    MultiRunner& R = multi_runners()[dev];
    std::lock_guard<std::mutex> lk(R.m);
    R.graphs = on != 0;
    return CM_OK;
}

int cm_multi_stats(int64_t* rounds, int64_t* launches)
{
    if (!rounds || !launches) return fail(CM_EINVAL, "null argument");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MR_MAX_DEVICES) return fail(CM_EUNSUPPORTED, "device index beyond the supported range");
    MultiRunner& R = multi_runners()[dev];
    std::lock_guard<std::mutex> lk(R.m);
    *rounds = R.rounds;
    *launches = R.launches;
    return CM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// entry point that allocates nothing (cm_mddf_noalloc)
// ---------------------------------------------------------------------------
// Fixed device scratch: every device buffer of such a call (integer hit
// counters, group maps, the batch buffers of the cell lists, placements and
// sorted molecules) is carved from it; the caller brings the coordinates and
// the FP64 output.  One copy per device, held until the process ends.
constexpr size_t SCRATCH_BYTES = (size_t)512 << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

constexpr int NA_MAX_DEVICES = 64;
// room for the 256-byte alignment of the carved buffers and the few bytes
// per frame (chunk offsets, drawn molecules) the per-frame estimate leaves out
constexpr int64_t NA_SLACK = 65536;

struct NoAllocDevice {
    std::mutex m;          // serialises the calls on this device
    bool ready = false;
    void* ptr = nullptr;   // device address of `scratch` on this device
    cm_job* job = nullptr; // host-side state reused by every call
};

NoAllocDevice* na_devices()
{
    static NoAllocDevice* d = new NoAllocDevice[NA_MAX_DEVICES];
    return d;
}

// device bytes a call needs: the integer counters and fixed buffers, plus
// the batch buffers of one frame
int64_t noalloc_fixed_bytes(const cm_job* J)
{
    int64_t b = 8 * J->acc_len + 8 * 4 * (int64_t)J->d.nbins + 64 + 4 * 256;
    if (J->d.groups_u_ptr) b += 4 * ((int64_t)J->d.natoms_per_mol_u + 2 + J->d.groups_u_ptr[J->d.natoms_per_mol_u]);
    if (J->d.groups_v_ptr) b += 4 * (J->Nv + 2 + J->d.groups_v_ptr[J->Nv]);
    return b;
}

}  // namespace

extern "C" {

int cm_mddf_noalloc(const cm_job_desc* desc, int32_t nframes, const int64_t* frame_index, const void* d_x_solute,
                    const void* d_x_solvent, int32_t dtype, const double* cells, const double* weights, double* d_acc,
                    double* volume_total, void* stream)
{
    if (!desc || !d_acc || !volume_total) return fail(CM_EINVAL, "null argument");
    if (nframes < 1 || !frame_index || !d_x_solute || !d_x_solvent || !cells || !weights)
        return fail(CM_EINVAL, "invalid frame arrays");
    if (dtype != CM_F64 && dtype != CM_F32) return fail(CM_EINVAL, "dtype must be CM_F64 or CM_F32");
    if (int r = validate_desc(*desc)) return r;
    for (int f = 0; f < nframes; f++) {
        if (!(weights[f] > 0.0) || !std::isfinite(weights[f])) return fail(CM_EINVAL, "frame weights must be positive and finite");
        if (frame_index[f] < 1 || frame_index[f] > 0xFFFFFFFFll) return fail(CM_EINVAL, "frame_index must be in [1, 2^32)");
    }
    // sizes first, on the host: a job that does not fit is refused before
    // any device work
    const int ds = dtype == CM_F32 ? 4 : 8;
    int64_t need_fixed, capmax, kcapmax, Fmax, pfb;
    {
        cm_job plan;
        job_params(&plan, *desc);
        need_fixed = noalloc_fixed_bytes(&plan);
        plan.d.memory_budget = (int64_t)SCRATCH_BYTES;
        if (int r = batch_plan(&plan, nframes, cells, ds, false, &capmax, &kcapmax, &Fmax, &pfb)) return r;
        if (need_fixed + pfb + NA_SLACK > (int64_t)SCRATCH_BYTES)
            return fail(CM_ENOMEM, "job too large for the no-allocation entry point: it needs " +
                                       std::to_string(need_fixed + pfb) + " bytes of device scratch for one frame, the "
                                       "fixed scratch holds " + std::to_string(SCRATCH_BYTES));
        Fmax = std::min<int64_t>(Fmax, std::max<int64_t>(1, ((int64_t)SCRATCH_BYTES - need_fixed - NA_SLACK) / pfb));
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(CM_EUNSUPPORTED, "device index beyond the supported range");
    NoAllocDevice& D = na_devices()[dev];
    std::lock_guard<std::mutex> lk(D.m);
    if (!D.ready) {
        // once per device, with that device current: the device address of
        // this device's copy of the scratch array
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(CM_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        D.ptr = p;
        D.job = new cm_job();
        D.ready = true;
    }
    cm_job* J = D.job;
    // every buffer of the previous call is given up: the arena starts over
    DevBuf* bufs[] = {&J->cnt, &J->acc, &J->hits, &J->stats, &J->gmap, &J->geo, &J->cells_cnt, &J->cells_start,
                      &J->tmp4, &J->tmpcell, &J->sorted, &J->solw, &J->ucen, &J->vw, &J->vcell, &J->vcen, &J->P,
                      &J->vw_r, &J->vcen_r, &J->bflag, &J->blist, &J->nbulk, &J->uk, &J->drawn_ptr,
                      &J->drawn_idx, &J->xin_u, &J->xin_v, &J->near, &J->mkey, &J->mrank, &J->mcnt, &J->mstart,
                      &J->msorted, &J->nnear, &J->chunk_base};
    for (DevBuf* b : bufs) *b = DevBuf{};
    J->arena = (char*)D.ptr;
    J->arena_cap = SCRATCH_BYTES;
    J->arena_used = 0;
    J->stream = (hipStream_t)stream;
    // the caller's output is the FP64 accumulator block (zeroed by the setup)
    const int64_t acc_len = 4 * (int64_t)desc->nbins + 2 * (int64_t)desc->n_groups_u * desc->nbins +
                            2 * (int64_t)desc->n_groups_v * desc->nbins;
    J->acc = DevBuf{d_acc, 8 * (size_t)acc_len};
    if (int r = job_setup(J, *desc)) return r;
    J->d.max_batch_frames = Fmax;
    J->d.memory_budget = (int64_t)SCRATCH_BYTES;
    // the batch buffers for the largest batch, carved once
    if (int r = alloc_batch(J, (int)Fmax, capmax, kcapmax, ds, false)) return r;
    if (int r = submit_common(J, nframes, frame_index, d_x_solute, d_x_solvent, dtype, cells, weights, false)) return r;
    if (J->pending) {
        if (int r = flush(J, J->pending_w)) return r;
        J->pending = false;
    }
    uint64_t gerr = 0;
    HIPCHK(hipMemcpyAsync(&gerr, (char*)J->stats.p + 16, 8, hipMemcpyDeviceToHost, J->stream));
    HIPCHK(hipStreamSynchronize(J->stream));
    J->finished = true;
    if (int r = geometry_error(gerr)) return r;
    *volume_total = J->volume_total;
    return CM_OK;
}

int cm_device_download(void* host, const void* dptr, size_t bytes)
{
    HIPCHK(hipMemcpy(host, dptr, bytes, hipMemcpyDeviceToHost));
    return CM_OK;
}

}  // extern "C"
