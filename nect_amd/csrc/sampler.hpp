// Ray geometry of the cone/parallel-beam sampler, device side.
//
// Semantics follow nect/sampling/ct_sampling.cu:117-576: per-view source and
// detector points in double, float view-angle cosine/sine, a double slab test
// against the box shrunk by a quarter voxel per side, float ray length, double
// normalised steps, float sample positions, break tolerance 1e-5.  The
// library is compiled with -ffp-contract=off so every operation below rounds
// once, as written.
//
// Design difference: the reference walks every sample of a ray in one thread
// and writes 12*S bytes per lane (uncoalesced).  Here a ray is reduced to a
// 64-byte descriptor (normalised start, step, count); the count is found by a
// binary search over the break predicate, which is monotone in the sample
// index (each coordinate is a monotone function of i, float rounding is
// monotone), and any consumer recomputes sample i as float(start + i*step).
#pragma once

#include "common.hpp"

namespace nect {

struct Pt {
    double x, y, z;
};

// Per-view constants, prepared once on the host (see make_geom_const).
struct GeomConst {
    NectGeometry g;
    float c, s;          // cos/sin of the view angle, (float)cos((double)angle)
    float R[6];          // R12 R13 R22 R23 R32 R33 of the detector rotation
    int32_t has_rot;
    float par_offset;    // sqrtf(sX*sX + sY*sY), parallel-beam source distance
    float hx, hy, hz;    // half extents of the shrunk box (float)
    float two_r;         // 2*radius (float)
    float z_lo, z_hi;    // cylindrical z window: 0 + remove_top, 1 - remove_bottom
    int32_t n_pixels;    // rows*cols
};

// Host: fold a geometry and a view angle into GeomConst (float ops only).
GeomConst make_geom_const(const NectGeometry &g, float angle_rad);

__device__ __forceinline__ double smax_d(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin_d(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double clamp01_fmaxmin(double v) { return fmin(fmax(v, 0.0), 1.0); }
// torch.clamp_(0, 1): NaN passes through.
__device__ __forceinline__ float clamp01_nan(float v) { return (v < 0.0f) ? 0.0f : ((v > 1.0f) ? 1.0f : v); }

__device__ __forceinline__ void rotate_view(Pt &a, Pt &b, const GeomConst &k) {
    const double c = (double)k.c, s = (double)k.s, cor = (double)k.g.cor;
    a.y = a.y - cor;
    b.y = b.y - cor;
    const double ax = c * a.x - s * a.y, ay = s * a.x + c * a.y;
    const double bx = c * b.x - s * b.y, by = s * b.x + c * b.y;
    a.x = ax; a.y = ay;
    b.x = bx; b.y = by;
}

// Detector pixel -> detector point in the unrotated frame (ct_sampling.cu:226-245).
__device__ __forceinline__ Pt detector_point(const GeomConst &k, int ray, float ou, float ov) {
    const int cols = k.g.n_det[1];
    const int col = ray % cols;
    const int row = ray / cols;
    const float half_c = (float)(cols - 1) / 2;
    const float half_r = (float)(k.g.n_det[0] - 1) / 2;
    const float u = (((float)col - half_c) + ou) * k.g.d_det[1] + k.g.off_det[1];
    const float v = (((float)row - half_r) + ov) * k.g.d_det[0] + k.g.off_det[0];
    Pt p = {0.0, -(double)u, (double)v};
    if (k.has_rot) {
        const float x = (float)((double)k.R[0] * p.y + (double)k.R[1] * p.z);
        const float y = (float)((double)k.R[2] * p.y + (double)k.R[3] * p.z);
        const float z = (float)((double)k.R[4] * p.y + (double)k.R[5] * p.z);
        p.x = x; p.y = y; p.z = z;
    }
    return p;
}

// Slab test (ct_sampling.cu:160-198): both parameters must lie in [0, 1].
__device__ __forceinline__ bool slab(const Pt &a, const Pt &b, const GeomConst &k, Pt &i1, Pt &i2) {
    const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    const double hx = (double)k.hx, hy = (double)k.hy, hz = (double)k.hz;
    double tx0 = (-hx - a.x) / dx, tx1 = (hx - a.x) / dx;
    if (tx0 > tx1) { const double t = tx0; tx0 = tx1; tx1 = t; }
    double ty0 = (-hy - a.y) / dy, ty1 = (hy - a.y) / dy;
    if (ty0 > ty1) { const double t = ty0; ty0 = ty1; ty1 = t; }
    double tz0 = (-hz - a.z) / dz, tz1 = (hz - a.z) / dz;
    if (tz0 > tz1) { const double t = tz0; tz0 = tz1; tz1 = t; }
    const double t0 = smax_d(smax_d(tx0, ty0), tz0);
    const double t1 = smin_d(smin_d(tx1, ty1), tz1);
    if (t0 > t1 || !(t0 >= 0.0 && t0 <= 1.0) || !(t1 >= 0.0 && t1 <= 1.0)) return false;
    i1.x = a.x + t0 * dx; i1.y = a.y + t0 * dy; i1.z = a.z + t0 * dz;
    i2.x = a.x + t1 * dx; i2.y = a.y + t1 * dy; i2.z = a.z + t1 * dz;
    return true;
}

// Ray/cylinder intersection (ct_sampling.cu:258-288), float quadratic.
__device__ __forceinline__ bool cylinder(const Pt &p1, const Pt &d, const GeomConst &k, Pt &st, Pt &en) {
    const float a = (float)(d.x * d.x + d.y * d.y);
    const float b = (float)(2 * p1.x * d.x + 2 * (p1.y - (double)k.g.cor) * d.y);
    const float rr = k.g.radius * k.g.radius;
    const float cc = (float)(p1.x * p1.x + p1.y * p1.y - (double)rr);
    const float disc = b * b - (4 * a) * cc;
    if (disc <= 0) return false;
    const float sq = sqrtf(disc);
    const float two_a = 2 * a;
    const float t1 = (-b - sq) / two_a;
    const float t2 = (-b + sq) / two_a;
    st.x = p1.x + (double)t1 * d.x; st.y = p1.y + (double)t1 * d.y; st.z = (double)t1 * d.z;
    en.x = p1.x + (double)t2 * d.x; en.y = p1.y + (double)t2 * d.y; en.z = (double)t2 * d.z;
    rotate_view(st, en, k);
    return true;
}

// Entry/exit points in millimetres (ct_sampling.cu:290-389).
__device__ __forceinline__ bool entry_exit(const GeomConst &k, int ray, float ou, float ov, Pt &st, Pt &en) {
    Pt p2 = detector_point(k, ray, ou, ov);
    const int kind = k.g.kind;
    if (kind == NECT_CONE_VOXEL || kind == NECT_PARALLEL_VOXEL) {
        Pt p1;
        if (kind == NECT_CONE_VOXEL) {
            p2.x = p2.x + (double)k.g.dod;
            p1.x = -(double)k.g.dso; p1.y = 0.0; p1.z = 0.0;
        } else {
            p2.x = p2.x + (double)k.par_offset;
            p1.x = -p2.x; p1.y = p2.y; p1.z = p2.z;
        }
        rotate_view(p1, p2, k);
        return slab(p1, p2, k, st, en);
    }
    if (kind == NECT_CONE_CYLINDRICAL) {
        const Pt p1 = {-(double)k.g.dso, 0.0, 0.0};
        p2.x = p2.x + (double)k.g.dod;
        const Pt d = {p2.x - p1.x, p2.y, p2.z};
        return cylinder(p1, d, k, st, en);
    }
    p2.x = p2.x + (double)k.g.radius;
    const Pt p1 = {-p2.x, p2.y, p2.z};
    const Pt d = {p2.x - p1.x, 0.0, 0.0};
    return cylinder(p1, d, k, st, en);
}

__device__ __forceinline__ float coord(double start, double step, int i) {
    return (float)(start + (double)i * step);
}

// The reference loop's break test at sample i (ct_sampling.cu:537-571 voxel,
// :440-467 cylindrical).
__device__ __forceinline__ bool stops_at(const GeomConst &k, const Pt &st, const Pt &en, const Pt &sp, int i) {
    const float px = coord(st.x, sp.x, i), py = coord(st.y, sp.y, i), pz = coord(st.z, sp.z, i);
    if (k.g.kind == NECT_CONE_VOXEL || k.g.kind == NECT_PARALLEL_VOXEL) {
        const float eps = 1e-5f;
        bool b = false;
        if (sp.x > 0) b |= (double)(px - eps) > en.x;
        else if (sp.x < 0) b |= (double)(px + eps) < en.x;
        if (sp.y > 0) b |= (double)(py - eps) > en.y;
        else if (sp.y < 0) b |= (double)(py + eps) < en.y;
        if (sp.z > 0) b |= (double)(pz - eps) > en.z;
        else if (sp.z < 0) b |= (double)(pz + eps) < en.z;
        return b;
    }
    bool b = false;
    if (sp.x > 0) b = (double)px > en.x;
    else if (sp.x < 0) b = (double)px < en.x;
    else if (sp.y > 0) b = (double)py > en.y;
    else if (sp.y < 0) b = (double)py < en.y;
    return b || pz < k.z_lo || pz > k.z_hi;
}

// Descriptor of one ray (ct_sampling.cu:392-576 without the per-sample writes).
__device__ __forceinline__ void describe_ray(const GeomConst &k, int ray, float ou, float ov, float jitter, int S,
                                             float step_mm, int uniform, NectRay &out) {
    out.count = 0;
    out.distance = 0.0f;
    out.start[0] = out.start[1] = out.start[2] = 0.0;
    out.step[0] = out.step[1] = out.step[2] = 0.0;
    Pt st, en;
    if (!entry_exit(k, ray, ou, ov, st, en)) return;
    Pt sp = {en.x - st.x, en.y - st.y, en.z - st.z};
    const float rd = (float)sqrt(sp.x * sp.x + sp.y * sp.y + sp.z * sp.z);
    const bool voxel = (k.g.kind == NECT_CONE_VOXEL || k.g.kind == NECT_PARALLEL_VOXEL);
    double nx, ny, nz;
    if (voxel) {
        nx = (double)k.g.s_voxel[2]; ny = (double)k.g.s_voxel[1]; nz = (double)k.g.s_voxel[0];
    } else {
        nx = (double)k.two_r; ny = (double)k.two_r; nz = (double)k.g.height;
    }
    if (uniform) {
        out.distance = step_mm;
        sp.x = (sp.x / (double)rd) * (double)step_mm / nx;
        sp.y = (sp.y / (double)rd) * (double)step_mm / ny;
        sp.z = (sp.z / (double)rd) * (double)step_mm / nz;
    } else {
        out.distance = rd / (float)S;
        sp.x = (sp.x / (double)S) / nx;
        sp.y = (sp.y / (double)S) / ny;
        sp.z = (sp.z / (double)S) / nz;
    }
    st.x /= nx; en.x /= nx;
    st.y /= ny; en.y /= ny;
    st.z /= nz;
    if (voxel) en.z /= nz;
    const double j = (double)jitter;
    st.x = (st.x + 0.5) + j * sp.x;
    st.y = (st.y + 0.5) + j * sp.y;
    st.z = (st.z + 0.5) + j * sp.z;
    if (voxel) {
        en.x = (en.x + 0.5) + sp.x;
        en.y = (en.y + 0.5) + sp.y;
        en.z = (en.z + 0.5) + sp.z;
    } else {
        en.x = en.x + 0.5;
        en.y = en.y + 0.5;
    }
    st.x = clamp01_fmaxmin(st.x); st.y = clamp01_fmaxmin(st.y); st.z = clamp01_fmaxmin(st.z);
    en.x = clamp01_fmaxmin(en.x); en.y = clamp01_fmaxmin(en.y); en.z = clamp01_fmaxmin(en.z);
    // First index whose break test fires (S if none): binary search over the
    // monotone predicate.
    int count;
    if (stops_at(k, st, en, sp, 0)) {
        count = 0;
    } else {
        int lo = 0, hi = S;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (stops_at(k, st, en, sp, mid)) hi = mid; else lo = mid;
        }
        count = hi;
    }
    out.count = count;
    out.start[0] = st.x; out.start[1] = st.y; out.start[2] = st.z;
    out.step[0] = sp.x; out.step[1] = sp.y; out.step[2] = sp.z;
}

// Sample i of a described ray, as the trainer sees it after clamp_(0, 1):
// (z, y, x) in p[0..2]; returns whether the reference trainer keeps it
// (i < count and not the all-zero row masked at base_trainer.py:523-528).
__device__ __forceinline__ bool ray_sample(const NectRay &r, int i, float p[3]) {
    p[0] = clamp01_nan(coord(r.start[2], r.step[2], i));
    p[1] = clamp01_nan(coord(r.start[1], r.step[1], i));
    p[2] = clamp01_nan(coord(r.start[0], r.step[0], i));
    return i < r.count && !(p[0] == 0.0f && p[1] == 0.0f && p[2] == 0.0f);
}

}  // namespace nect
