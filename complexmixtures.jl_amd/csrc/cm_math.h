// cm_math.h -- host/device arithmetic of the MDDF path whose bits must be
// reproducible: the counter-based random stream, sin/cos of 2*pi*u, the
// minimum-image distance, the rigid-body placement and the histogram bin.
//
// Everything here is compiled with -ffp-contract=off: expressions evaluate as
// written, with fused multiply-adds only where fma() is spelled out, so the
// host build, the device build and any CPU checker that follows the same
// definitions agree to the last bit.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define CM_HD __host__ __device__ __forceinline__

namespace cm {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Replaces the Xoroshiro128+/StableRNG streams of
// src/rigid_body.jl:6-19, which cannot be reproduced outside Julia.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

CM_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c.z;
        u32x4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Fourth counter word: which stream a draw belongs to.
constexpr uint32_t TAG_REF = 0x52454631u;    // reference-solute draws of a frame
constexpr uint32_t TAG_SLOT = 0x534c4f54u;   // placement draws of one solvent slot (+1: second block)

CM_HD double u01(uint32_t r) { return (double)r * 0x1p-32; }
// Uniform index in [0, n): multiply-shift of a 32-bit draw.
CM_HD int64_t draw_index(uint32_t r, int64_t n) { return (int64_t)(((uint64_t)r * (uint64_t)n) >> 32); }

// ---------------------------------------------------------------------------
// sin(2*pi*u), cos(2*pi*u): exact quarter-turn reduction (r = u - q/4 is exact
// for u in [0,1)), then the Taylor series of sin(2*pi*r) and cos(2*pi*r) for
// |r| <= 1/8 in Horner form.  Coefficients are (-1)^k (2pi)^n / n! rounded to
// double; the first omitted terms are below 1e-19.
// ---------------------------------------------------------------------------
CM_HD void sincos2pi(double u, double* s_out, double* c_out)
{
    const double S0 = 0x1.921fb54442d18p+2, S1 = -0x1.4abbce625be53p+5, S2 = 0x1.466bc6775aae2p+6,
                 S3 = -0x1.32d2cce62bd86p+6, S4 = 0x1.50783487ee782p+5, S5 = -0x1.e3074fde8871fp+3,
                 S6 = 0x1.e8f434d018d63p+1, S7 = -0x1.6fadb9f155744p-1, S8 = 0x1.aaec32af93359p-4;
    const double C0 = 0x1.0000000000000p+0, C1 = -0x1.3bd3cc9be45dep+4, C2 = 0x1.03c1f081b5ac4p+6,
                 C3 = -0x1.55d3c7e3cbffap+6, C4 = 0x1.e1f506891babbp+5, C5 = -0x1.a6d1f2a204a8cp+4,
                 C6 = 0x1.f9d38a3763cc3p+2, C7 = -0x1.b6e24f44b128fp+0, C8 = 0x1.20c62c2f2d7f5p-2,
                 C9 = -0x1.2a0c591af8314p-5;
    double q = rint(4.0 * u);
    double r = (4.0 * u - q) * 0.25;
    double z = r * r;
    double ps = S8;
    ps = fma(ps, z, S7); ps = fma(ps, z, S6); ps = fma(ps, z, S5); ps = fma(ps, z, S4);
    ps = fma(ps, z, S3); ps = fma(ps, z, S2); ps = fma(ps, z, S1); ps = fma(ps, z, S0);
    double s = r * ps;
    double pc = C9;
    pc = fma(pc, z, C8); pc = fma(pc, z, C7); pc = fma(pc, z, C6); pc = fma(pc, z, C5);
    pc = fma(pc, z, C4); pc = fma(pc, z, C3); pc = fma(pc, z, C2); pc = fma(pc, z, C1);
    pc = fma(pc, z, C0);
    double c = pc;
    int qm = ((int)q) & 3;
    if (qm == 0) { *s_out = s; *c_out = c; }
    else if (qm == 1) { *s_out = c; *c_out = -s; }
    else if (qm == 2) { *s_out = -s; *c_out = -c; }
    else { *s_out = -c; *c_out = s; }
}

// eulermat(beta, gamma, theta) of src/rigid_body.jl:45-57, angles given in
// turns (angle = 2*pi*u).  Row-major output.
CM_HD void eulermat_turns(double ub, double ug, double ut, double A[9])
{
    double s1, c1, s2, c2, s3, c3;
    sincos2pi(ub, &s1, &c1);
    sincos2pi(ug, &s2, &c2);
    sincos2pi(ut, &s3, &c3);
    A[0] = c2 * c3;                A[1] = -c2 * s3;               A[2] = s2;
    A[3] = c1 * s3 + c3 * s1 * s2; A[4] = c1 * c3 - s1 * s2 * s3; A[5] = -c2 * s1;
    A[6] = s1 * s3 - c1 * c3 * s2; A[7] = c1 * s2 * s3 + c3 * s1; A[8] = c1 * c2;
}

// ---------------------------------------------------------------------------
// Unit cell.  H holds the lattice vectors as columns (column major, H[3c+r]),
// the convention of the reference's unit cells (src/Trajectory.jl:72-81).
// ---------------------------------------------------------------------------
struct Cell64 { double H[9]; double Hi[9]; };

CM_HD void cell_setup(const double* cell9, int triclinic, Cell64* c)
{
    for (int k = 0; k < 9; k++) c->H[k] = cell9[k];
    if (!triclinic) {  // convert_unitcell(::SVector, m) keeps the diagonal
        c->H[1] = 0.0; c->H[2] = 0.0; c->H[3] = 0.0; c->H[5] = 0.0; c->H[6] = 0.0; c->H[7] = 0.0;
    }
    const double* H = c->H;
    double m00 = H[0], m10 = H[1], m20 = H[2], m01 = H[3], m11 = H[4], m21 = H[5], m02 = H[6], m12 = H[7], m22 = H[8];
    double c00 = m11 * m22 - m12 * m21;
    double c01 = m12 * m20 - m10 * m22;
    double c02 = m10 * m21 - m11 * m20;
    double det = m00 * c00 + m01 * c01 + m02 * c02;
    double c10 = m02 * m21 - m01 * m22;
    double c11 = m00 * m22 - m02 * m20;
    double c12 = m01 * m20 - m00 * m21;
    double c20 = m01 * m12 - m02 * m11;
    double c21 = m02 * m10 - m00 * m12;
    double c22 = m00 * m11 - m01 * m10;
    c->Hi[0] = c00 / det; c->Hi[1] = c01 / det; c->Hi[2] = c02 / det;
    c->Hi[3] = c10 / det; c->Hi[4] = c11 / det; c->Hi[5] = c12 / det;
    c->Hi[6] = c20 / det; c->Hi[7] = c21 / det; c->Hi[8] = c22 / det;
}

// cell_volume of src/mddf.jl:350-352: dot(cross(a, b), c)
CM_HD double cell_volume(const double* H)
{
    const double *a = H, *b = H + 3, *c = H + 6;
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    return x * c[0] + y * c[1] + z * c[2];
}

// Minimum image of a displacement: n = rint(Hinv dr), dr' = dr - H n.  Exact
// minimum image for every displacement shorter than half the smallest plane
// spacing of the cell.
CM_HD void minimage(const Cell64& c, double dx, double dy, double dz, double* ox, double* oy, double* oz)
{
    const double *Hi = c.Hi, *H = c.H;
    double fx = fma(Hi[0], dx, fma(Hi[3], dy, Hi[6] * dz));
    double fy = fma(Hi[1], dx, fma(Hi[4], dy, Hi[7] * dz));
    double fz = fma(Hi[2], dx, fma(Hi[5], dy, Hi[8] * dz));
    double nx = rint(fx), ny = rint(fy), nz = rint(fz);
    *ox = fma(-H[0], nx, fma(-H[3], ny, fma(-H[6], nz, dx)));
    *oy = fma(-H[1], nx, fma(-H[4], ny, fma(-H[7], nz, dy)));
    *oz = fma(-H[2], nx, fma(-H[5], ny, fma(-H[8], nz, dz)));
}

// The distance that is recorded: squared minimum-image distance in FP64.
CM_HD double canon_d2(const Cell64& c, double x0, double x1, double x2, double y0, double y1, double y2)
{
    double ox, oy, oz;
    minimage(c, y0 - x0, y1 - x1, y2 - x2, &ox, &oy, &oz);
    return fma(ox, ox, fma(oy, oy, oz * oz));
}

// setbin of src/results.jl:28, 1-based
CM_HD int64_t setbin(double d, double step)
{
    double b = ceil(d / step);
    int64_t ib = (int64_t)b;
    return ib < 1 ? 1 : ib;
}

// random_move! of src/rigid_body.jl:107-137 on one molecule: wrap relative
// to the reference atom, rotate by eulermat about the centre of coordinates
// (move!, src/rigid_body.jl:73-80), translate the centre to H*uc.  x(a, k)
// reads the source atoms, out(a, x, y, z) receives the placed ones; the
// wrapped positions are formed twice (for the centre, then for the
// placement) with the same operations, so nothing intermediate is stored.
// In place (out writing where x reads) is allowed: atom a is read before it
// is written and the reference atom is read first.
template <class Src, class Dst>
CM_HD void random_move(const Src& x, Dst& out, int n, int iref, const Cell64& c, const double uc[3],
                       const double ua[3])
{
    const double rx = x(iref, 0), ry = x(iref, 1), rz = x(iref, 2);
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int a = 0; a < n; a++) {
        double ox, oy, oz;
        minimage(c, x(a, 0) - rx, x(a, 1) - ry, x(a, 2) - rz, &ox, &oy, &oz);
        cx += rx + ox; cy += ry + oy; cz += rz + oz;
    }
    cx = cx / (double)n; cy = cy / (double)n; cz = cz / (double)n;
    const double* H = c.H;
    const double ncm0 = H[0] * uc[0] + H[3] * uc[1] + H[6] * uc[2];
    const double ncm1 = H[1] * uc[0] + H[4] * uc[1] + H[7] * uc[2];
    const double ncm2 = H[2] * uc[0] + H[5] * uc[1] + H[8] * uc[2];
    double A[9];
    eulermat_turns(ua[0], ua[1], ua[2], A);
    for (int a = 0; a < n; a++) {
        double ox, oy, oz;
        minimage(c, x(a, 0) - rx, x(a, 1) - ry, x(a, 2) - rz, &ox, &oy, &oz);
        const double dx = (rx + ox) - cx, dy = (ry + oy) - cy, dz = (rz + oz) - cz;
        out(a, A[0] * dx + A[1] * dy + A[2] * dz + ncm0,
               A[3] * dx + A[4] * dy + A[5] * dz + ncm1,
               A[6] * dx + A[7] * dy + A[8] * dz + ncm2);
    }
}

}  // namespace cm
