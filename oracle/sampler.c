/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this file's library, and only as the
 * checker; the product never links or calls it.
 *
 * Plain-C restatement of NeCT's ray sampler, following
 * nect/sampling/ct_sampling.cu operation by operation.  Every arithmetic step
 * keeps the precision the reference gives it (double `Point3D` coordinates,
 * float geometry fields, float ray length, double steps, float sample
 * positions).  Built with -ffp-contract=off so each operation rounds once,
 * exactly as written.
 *
 * Fixed points of the specification shared with the GPU library:
 *   - the view-angle cosine/sine are `(float)cos((double)angle)` and
 *     `(float)sin((double)angle)` (the reference calls CUDA cosf/sinf on the
 *     device, which cannot be reproduced; ct_sampling.cu:122-123);
 *   - the detector-rotation matrix entries are the float products of
 *     `(float)cos(-(double)a)` etc. (ct_sampling.cu:200-224);
 *   - the parallel-beam source offset `sqrtf(sX*sX + sY*sY)` is evaluated in
 *     float (ct_sampling.cu:370).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

enum { ORC_CONE_VOXEL = 0, ORC_PARALLEL_VOXEL = 1, ORC_CONE_CYL = 2, ORC_PARALLEL_CYL = 3 };

/* Field-for-field the geometry structs of ct_sampling.cu:7-115 (all float). */
typedef struct {
    int32_t kind;
    int32_t n_det[2];     /* rows, cols */
    float d_det[2];
    float off_det[2];
    float rot_det[3];     /* roll, pitch, yaw */
    float cor;
    float dso, dod;
    float s_voxel[3];     /* z, y, x */
    float d_voxel[3];
    float radius, height, remove_top, remove_bottom;
} orc_geometry;

typedef struct { double x, y, z; } pt3;

/* std::max / std::min as the reference's device code evaluates them. */
static double smax(double a, double b) { return (a < b) ? b : a; }
static double smin(double a, double b) { return (b < a) ? b : a; }
/* CUDA's max(double,float)/min(double,float) overloads are fmax/fmin. */
static double clamp01_fm(double v) { return fmin(fmax(v, 0.0), 1.0); }

/* ct_sampling.cu:121-128 */
static void rotate_points(pt3 *s, pt3 *e, float c, float sn, float cor) {
    s->y = s->y - (double)cor;
    e->y = e->y - (double)cor;
    pt3 ns = {(double)c * s->x - (double)sn * s->y, (double)sn * s->x + (double)c * s->y, s->z};
    pt3 ne = {(double)c * e->x - (double)sn * e->y, (double)sn * e->x + (double)c * e->y, e->z};
    *s = ns;
    *e = ne;
}

/* ct_sampling.cu:160-198 */
static int find_intersection(pt3 a, pt3 b, pt3 mn, pt3 mx, pt3 *i1, pt3 *i2) {
    double tMinX = (mn.x - a.x) / (b.x - a.x);
    double tMaxX = (mx.x - a.x) / (b.x - a.x);
    if (tMinX > tMaxX) { double t = tMinX; tMinX = tMaxX; tMaxX = t; }
    double tMinY = (mn.y - a.y) / (b.y - a.y);
    double tMaxY = (mx.y - a.y) / (b.y - a.y);
    if (tMinY > tMaxY) { double t = tMinY; tMinY = tMaxY; tMaxY = t; }
    double tMinZ = (mn.z - a.z) / (b.z - a.z);
    double tMaxZ = (mx.z - a.z) / (b.z - a.z);
    if (tMinZ > tMaxZ) { double t = tMinZ; tMinZ = tMaxZ; tMaxZ = t; }
    double tMin = smax(smax(tMinX, tMinY), tMinZ);
    double tMax = smin(smin(tMaxX, tMaxY), tMaxZ);
    int in0 = (tMin >= 0.0 && tMin <= 1.0);
    int in1 = (tMax >= 0.0 && tMax <= 1.0);
    if (tMin > tMax || !in0 || !in1) return 0;
    i1->x = a.x + tMin * (b.x - a.x);
    i1->y = a.y + tMin * (b.y - a.y);
    i1->z = a.z + tMin * (b.z - a.z);
    i2->x = a.x + tMax * (b.x - a.x);
    i2->y = a.y + tMax * (b.y - a.y);
    i2->z = a.z + tMax * (b.z - a.z);
    return 1;
}

/* ct_sampling.cu:200-224: only the six entries the reference uses. */
static void detector_rotation(const orc_geometry *g, float R[6]) {
    float alpha = g->rot_det[2], beta = g->rot_det[1], gamma = g->rot_det[0];
    float ca = (float)cos(-(double)alpha), cb = (float)cos(-(double)beta), cg = (float)cos(-(double)gamma);
    float sa = (float)sin(-(double)alpha), sb = (float)sin(-(double)beta), sg = (float)sin(-(double)gamma);
    R[0] = ca * sb * sg - sa * cg; /* R12 */
    R[1] = ca * sb * cg + sa * sg; /* R13 */
    R[2] = sa * sb * sg + ca * cg; /* R22 */
    R[3] = sa * sb * cg - ca * sg; /* R23 */
    R[4] = cb * sg;                /* R32 */
    R[5] = cb * cg;                /* R33 */
}

/* ct_sampling.cu:226-245 */
static pt3 detector_pos(const orc_geometry *g, int ray_index, float off_u, float off_v) {
    int col = ray_index % g->n_det[1];
    int row = ray_index / g->n_det[1];
    float half_c = (float)(g->n_det[1] - 1) / 2;
    float half_r = (float)(g->n_det[0] - 1) / 2;
    float u = (((float)col - half_c) + off_u) * g->d_det[1] + g->off_det[1];
    float v = (((float)row - half_r) + off_v) * g->d_det[0] + g->off_det[0];
    pt3 p = {0.0, -(double)u, (double)v};
    if (g->rot_det[0] != 0 || g->rot_det[1] != 0 || g->rot_det[2] != 0) {
        float R[6];
        detector_rotation(g, R);
        float x = (float)((double)R[0] * p.y + (double)R[1] * p.z);
        float y = (float)((double)R[2] * p.y + (double)R[3] * p.z);
        float z = (float)((double)R[4] * p.y + (double)R[5] * p.z);
        p.x = x; p.y = y; p.z = z;
    }
    return p;
}

/* ct_sampling.cu:258-288 */
static int cyl_helper(pt3 p1, pt3 d, const orc_geometry *g, float c, float s, pt3 *st, pt3 *en) {
    float a = (float)(d.x * d.x + d.y * d.y);
    float b = (float)(2 * p1.x * d.x + 2 * (p1.y - (double)g->cor) * d.y);
    float rr = g->radius * g->radius;
    float cc = (float)(p1.x * p1.x + p1.y * p1.y - (double)rr);
    float disc = b * b - (4 * a) * cc;
    if (disc <= 0) return 0;
    float sq = sqrtf(disc);
    float two_a = 2 * a;
    float t1 = (-b - sq) / two_a;
    float t2 = (-b + sq) / two_a;
    st->x = p1.x + (double)t1 * d.x;
    st->y = p1.y + (double)t1 * d.y;
    st->z = (double)t1 * d.z;
    en->x = p1.x + (double)t2 * d.x;
    en->y = p1.y + (double)t2 * d.y;
    en->z = (double)t2 * d.z;
    rotate_points(st, en, c, s, g->cor);
    return 1;
}

/* ct_sampling.cu:290-389: entry/exit points of one ray, or 0 on a miss. */
static int start_end(const orc_geometry *g, int ray_index, float off_u, float off_v, float c, float s,
                     float par_offset, pt3 *st, pt3 *en) {
    pt3 p2 = detector_pos(g, ray_index, off_u, off_v);
    if (g->kind == ORC_CONE_VOXEL || g->kind == ORC_PARALLEL_VOXEL) {
        pt3 p1;
        if (g->kind == ORC_CONE_VOXEL) {
            p2.x = p2.x + (double)g->dod;
            p1.x = -(double)g->dso; p1.y = 0.0; p1.z = 0.0;
        } else {
            p2.x = p2.x + (double)par_offset;
            p1.x = -p2.x; p1.y = p2.y; p1.z = p2.z;
        }
        rotate_points(&p1, &p2, c, s, g->cor);
        float sX = (g->s_voxel[2] - g->d_voxel[2] / 2) / 2;
        float sY = (g->s_voxel[1] - g->d_voxel[1] / 2) / 2;
        float sZ = (g->s_voxel[0] - g->d_voxel[0] / 2) / 2;
        pt3 mn = {-(double)sX, -(double)sY, -(double)sZ};
        pt3 mx = {(double)sX, (double)sY, (double)sZ};
        return find_intersection(p1, p2, mn, mx, st, en);
    }
    if (g->kind == ORC_CONE_CYL) {
        pt3 p1 = {-(double)g->dso, 0.0, 0.0};
        p2.x = p2.x + (double)g->dod;
        pt3 d = {p2.x - p1.x, p2.y, p2.z};
        return cyl_helper(p1, d, g, c, s, st, en);
    }
    /* parallel cylindrical */
    p2.x = p2.x + (double)g->radius;
    pt3 p1 = {-p2.x, p2.y, p2.z};
    pt3 d = {p2.x - p1.x, 0.0, 0.0};
    return cyl_helper(p1, d, g, c, s, st, en);
}

/* ct_sampling.cu:475-576 (voxel) and :392-473 (cylindrical).  Writes the
 * (z,y,x) float samples into pts (row of S*3), returns the number written. */
static int all_points(const orc_geometry *g, pt3 st, pt3 en, float jitter, int S, float step_mm, int uniform,
                      float *dist, float *pts) {
    pt3 step = {en.x - st.x, en.y - st.y, en.z - st.z};
    float rd = (float)sqrt(step.x * step.x + step.y * step.y + step.z * step.z);
    int n = 0;
    if (g->kind == ORC_CONE_VOXEL || g->kind == ORC_PARALLEL_VOXEL) {
        double sx = g->s_voxel[2], sy = g->s_voxel[1], sz = g->s_voxel[0];
        if (uniform) {
            *dist = step_mm;
            step.x = (step.x / (double)rd) * (double)step_mm / sx;
            step.y = (step.y / (double)rd) * (double)step_mm / sy;
            step.z = (step.z / (double)rd) * (double)step_mm / sz;
        } else {
            *dist = rd / (float)S;
            step.x = (step.x / (double)S) / sx;
            step.y = (step.y / (double)S) / sy;
            step.z = (step.z / (double)S) / sz;
        }
        st.x /= sx; en.x /= sx;
        st.y /= sy; en.y /= sy;
        st.z /= sz; en.z /= sz;
        const float eps = 1e-5f;
        st.x = (st.x + 0.5) + (double)jitter * step.x;
        st.y = (st.y + 0.5) + (double)jitter * step.y;
        st.z = (st.z + 0.5) + (double)jitter * step.z;
        en.x = (en.x + 0.5) + step.x;
        en.y = (en.y + 0.5) + step.y;
        en.z = (en.z + 0.5) + step.z;
        st.z = clamp01_fm(st.z); st.y = clamp01_fm(st.y); st.x = clamp01_fm(st.x);
        en.z = clamp01_fm(en.z); en.y = clamp01_fm(en.y); en.x = clamp01_fm(en.x);
        for (int i = 0; i < S; i++) {
            float px = (float)(st.x + (double)i * step.x);
            float py = (float)(st.y + (double)i * step.y);
            float pz = (float)(st.z + (double)i * step.z);
            if (step.x > 0) { if ((double)(px - eps) > en.x) break; }
            else if (step.x < 0) { if ((double)(px + eps) < en.x) break; }
            if (step.y > 0) { if ((double)(py - eps) > en.y) break; }
            else if (step.y < 0) { if ((double)(py + eps) < en.y) break; }
            if (step.z > 0) { if ((double)(pz - eps) > en.z) break; }
            else if (step.z < 0) { if ((double)(pz + eps) < en.z) break; }
            pts[3 * i + 0] = pz;
            pts[3 * i + 1] = py;
            pts[3 * i + 2] = px;
            n = i + 1;
        }
        return n;
    }
    /* cylindrical */
    float two_r = 2 * g->radius;
    if (uniform) {
        *dist = step_mm;
        step.x = (step.x / (double)rd) * (double)step_mm / (double)two_r;
        step.y = (step.y / (double)rd) * (double)step_mm / (double)two_r;
        step.z = (step.z / (double)rd) * (double)step_mm / (double)g->height;
    } else {
        *dist = rd / (float)S;
        step.x = (step.x / (double)S) / (double)two_r;
        step.y = (step.y / (double)S) / (double)two_r;
        step.z = (step.z / (double)S) / (double)g->height;
    }
    st.x /= (double)two_r; en.x /= (double)two_r;
    st.y /= (double)two_r; en.y /= (double)two_r;
    st.z /= (double)g->height;
    st.x = (st.x + 0.5) + (double)jitter * step.x;
    st.y = (st.y + 0.5) + (double)jitter * step.y;
    st.z = (st.z + 0.5) + (double)jitter * step.z;
    en.x = en.x + 0.5;
    en.y = en.y + 0.5;
    st.z = clamp01_fm(st.z); st.y = clamp01_fm(st.y); st.x = clamp01_fm(st.x);
    en.z = clamp01_fm(en.z); en.y = clamp01_fm(en.y); en.x = clamp01_fm(en.x);
    float lo = 0 + g->remove_top;
    float hi = 1 - g->remove_bottom;
    for (int i = 0; i < S; i++) {
        float px = (float)(st.x + (double)i * step.x);
        float py = (float)(st.y + (double)i * step.y);
        float pz = (float)(st.z + (double)i * step.z);
        if (step.x > 0) { if ((double)px > en.x) break; }
        else if (step.x < 0) { if ((double)px < en.x) break; }
        else if (step.y > 0) { if ((double)py > en.y) break; }
        else if (step.y < 0) { if ((double)py < en.y) break; }
        if (pz < lo || pz > hi) break;
        pts[3 * i + 0] = pz;
        pts[3 * i + 1] = py;
        pts[3 * i + 2] = px;
        n = i + 1;
    }
    return n;
}

/* torch's in-place clamp_(0, 1) (projector.py:127): NaN propagates. */
static float clamp01_torch(float v) { return (v < 0.0f) ? 0.0f : ((v > 1.0f) ? 1.0f : v); }

/*
 * ct_sampling.cu:578-648 + projector.py:113-127 for a whole batch.
 * points: [num_rays, S, 3] (zero-filled here), distances: [num_rays],
 * counts: [num_rays] = samples written per ray (the reference leaves unwritten
 * slots at zero).  Points are clamped to [0,1] afterwards as the projector does.
 */
void oracle_sample(const orc_geometry *g, float angle_rad, const int64_t *ray_index, int64_t n_ray_index,
                   int32_t num_rays, int32_t starting_ray_index, int32_t S, float step_mm, int32_t uniform,
                   const float *jitter, const float *det_offset, float *points, float *distances,
                   int32_t *counts) {
    float c = (float)cos((double)angle_rad);
    float s = (float)sin((double)angle_rad);
    float par_offset = sqrtf(g->s_voxel[2] * g->s_voxel[2] + g->s_voxel[1] * g->s_voxel[1]);
    memset(points, 0, sizeof(float) * (size_t)num_rays * (size_t)S * 3);
    for (int32_t tid = 0; tid < num_rays; tid++) {
        distances[tid] = 0.0f;
        counts[tid] = 0;
        if ((int64_t)tid + starting_ray_index >= n_ray_index) continue;
        int ray = (int)ray_index[tid + starting_ray_index];
        if (ray >= g->n_det[0] * g->n_det[1]) continue;
        pt3 st, en;
        if (!start_end(g, ray, det_offset[2 * tid], det_offset[2 * tid + 1], c, s, par_offset, &st, &en)) continue;
        float *row = points + (size_t)tid * (size_t)S * 3;
        counts[tid] = all_points(g, st, en, jitter[tid], S, step_mm, uniform, &distances[tid], row);
        for (int k = 0; k < 3 * S; k++) row[k] = clamp01_torch(row[k]);
    }
}

/* Entry/exit points in millimetres for one ray (for the analytic checks). */
int oracle_ray_entry_exit(const orc_geometry *g, float angle_rad, int32_t ray, float off_u, float off_v,
                          double *out6) {
    float c = (float)cos((double)angle_rad);
    float s = (float)sin((double)angle_rad);
    float par_offset = sqrtf(g->s_voxel[2] * g->s_voxel[2] + g->s_voxel[1] * g->s_voxel[1]);
    pt3 st, en;
    if (!start_end(g, ray, off_u, off_v, c, s, par_offset, &st, &en)) return 0;
    out6[0] = st.x; out6[1] = st.y; out6[2] = st.z;
    out6[3] = en.x; out6[4] = en.y; out6[5] = en.z;
    return 1;
}
