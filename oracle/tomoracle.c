/*
 * ORACLE — CPU restatement of the TOM template-matching path.
 *
 * This file is TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it, and only as the checker; the
 * product library (tom_amd/libtomhip.so) never links or calls it.
 *
 * Contents (each cites the reference code it restates):
 *   orc_rotmat      rotation matrix of rot3d, incl. the exact cos/sin table
 *                   (Sptrans/tom_rotatec.c:146-291)
 *   orc_rot3d       trilinear inverse-mapped 3D rotation, float arithmetic in
 *                   the reference's evaluation order (Sptrans/tom_rotatec.c:293-398)
 *   orc_rot2d       bilinear 2D rotation of rot2d, float arithmetic in the
 *                   reference's evaluation order (Sptrans/tom_rotatec.c:49-142)
 *   orc_find_max    global max, first position in x-fastest order, strict '>'
 *                   (Analysis/tom_peakinc.c:23-52)
 *   orc_flcf_direct brute-force direct-space masked FLCF in FP64 for tiny
 *                   volumes (Oscar3/matlabPicker/search/tom_os3_corr.m:77-192,
 *                   statistics/tom_os3_mean.m:44-63, tom_os3_std.m:43-68)
 *
 * Pinning (see tests/test_oracle.py):
 *   orc_rotmat/orc_rot3d: bit-exact against the reference's own rot3dlin and
 *     tom_transf_init_rotmat_3d compiled from its sources (oracle/_ref) and the
 *     committed golden vectors; axis invariants of Sptrans/tom_rotate.m:59-61.
 *   orc_rot2d: identity at 0 degrees; a linear ramp rotates exactly (bilinear
 *     interpolation reproduces linear functions) with the reference's sense and
 *     pivot; out-of-image pixels are 0. On the last row (r_y == sy-1) rot2d
 *     takes no y tap (BB = 0), which leaves the value unchanged since vy2 = 0.
 *     Sptrans/tom_rotatec.c cannot be compiled here (its header defines MATLAB
 *     and includes mex.h), so no golden vectors exist for it.
 *   orc_find_max: numpy argmax (first occurrence) on x-fastest data; ties.
 *   orc_flcf_direct: the numpy FFT restatement agrees with it; planted copy
 *     scores exactly 1; affine invariance; zero local variance -> 0.
 *
 * Layout: every volume is x-fastest (MATLAB column-major), index
 * x + sx*(y + sy*z).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI ((double)3.14159265358979323846264338327950288419716939937510)

/* cos/sin of one Euler angle the way rot3d takes it: an exact table for the 16
 * "round" angles, else cos/sin in double of the float-rounded radian value. */
static void orc_cossin(float deg, float *c, float *s)
{
    static const float tab_deg[16] = {0, 30, 45, 60, 90, 120, 135, 150,
                                      180, 210, 225, 240, 270, 300, 315, 330};
    float tc[16], ts[16];
    tc[0] = 1;               ts[0] = 0;
    tc[1] = sqrt(3) / 2;     ts[1] = 0.5;
    tc[2] = sqrt(2) / 2;     ts[2] = sqrt(2) / 2;
    tc[3] = 0.5;             ts[3] = sqrt(3) / 2;
    tc[4] = 0;               ts[4] = 1;
    tc[5] = -0.5;            ts[5] = sqrt(3) / 2;
    tc[6] = -sqrt(2) / 2;    ts[6] = sqrt(2) / 2;
    tc[7] = -sqrt(3) / 2;    ts[7] = 0.5;
    tc[8] = -1;              ts[8] = 0;
    tc[9] = -sqrt(3) / 2;    ts[9] = -0.5;
    tc[10] = -sqrt(2) / 2;   ts[10] = -sqrt(2) / 2;
    tc[11] = -0.5;           ts[11] = -sqrt(3) / 2;
    tc[12] = 0;              ts[12] = -1;
    tc[13] = 0.5;            ts[13] = -sqrt(3) / 2;
    tc[14] = sqrt(2) / 2;    ts[14] = -sqrt(2) / 2;
    tc[15] = sqrt(3) / 2;    ts[15] = -0.5;
    int hit = 0;
    for (int i = 0; i < 16; i++) {
        if (tab_deg[i] == deg) { *c = tc[i]; *s = ts[i]; hit = 1; }
    }
    if (!hit) {
        float rad = deg * ORC_PI / 180; /* rounded to float, as rot3d does */
        *c = cos(rad);
        *s = sin(rad);
    }
}

/* rm[9] in the order rm00, rm10, rm20, rm01, rm11, rm21, rm02, rm12, rm22.
 * neuler == 3: [phi psi theta] in degrees; otherwise 9 matrix entries taken
 * in that same order (tom_rotatec.c:276-287). */
void orc_rotmat(const float *euler, int neuler, float *rm)
{
    if (neuler != 3) {
        for (int i = 0; i < 9; i++) rm[i] = euler[i];
        return;
    }
    float cphi, sphi, cpsi, spsi, cthe, sthe;
    orc_cossin(euler[0], &cphi, &sphi);
    orc_cossin(euler[1], &cpsi, &spsi);
    orc_cossin(euler[2], &cthe, &sthe);
    rm[0] = cpsi * cphi - cthe * spsi * sphi;   /* rm00 */
    rm[1] = spsi * cphi + cthe * cpsi * sphi;   /* rm10 */
    rm[2] = sthe * sphi;                        /* rm20 */
    rm[3] = -cpsi * sphi - cthe * spsi * cphi;  /* rm01 */
    rm[4] = -spsi * sphi + cthe * cpsi * cphi;  /* rm11 */
    rm[5] = sthe * cphi;                        /* rm21 */
    rm[6] = sthe * spsi;                        /* rm02 */
    rm[7] = -sthe * cpsi;                       /* rm12 */
    rm[8] = cthe;                               /* rm22 */
}

void orc_rot3d(const float *img, float *out, long sx, long sy, long sz,
               const float *euler, int neuler, float px, float py, float pz)
{
    float rm[9];
    orc_rotmat(euler, neuler, rm);
    const float rm00 = rm[0], rm10 = rm[1], rm20 = rm[2];
    const float rm01 = rm[3], rm11 = rm[4], rm21 = rm[5];
    const float rm02 = rm[6], rm12 = rm[7], rm22 = rm[8];
    const long sxy = sx * sy;
    for (long k = 0; k < sz; k++)
        for (long j = 0; j < sy; j++)
            for (long i = 0; i < sx; i++) {
                const long pi = (long)((float)i - px);
                const long pj = (long)((float)j - py);
                const long pk = (long)((float)k - pz);
                float *o = out + i + sx * (j + sy * k);
                float rx = px + rm00 * (float)pi + rm10 * (float)pj + rm20 * (float)pk;
                if (rx < 0 || rx > (float)(sx - 1)) { *o = 0; continue; }
                float ry = py + rm01 * (float)pi + rm11 * (float)pj + rm21 * (float)pk;
                if (ry < 0 || ry > (float)(sy - 1)) { *o = 0; continue; }
                float rz = pz + rm02 * (float)pi + rm12 * (float)pj + rm22 * (float)pk;
                if (rz < 0 || rz > (float)(sz - 1)) { *o = 0; continue; }
                const long fx = (long)rx, fy = (long)ry, fz = (long)rz;
                const float vx2 = rx - fx, vx1 = 1 - vx2;
                const float vy2 = ry - fy, vy1 = 1 - vy2;
                const float vz2 = rz - fz, vz1 = 1 - vz2;
                const long b = fx + fy * sx + fz * sxy;
                const int ex = (fx == sx - 1), ey = (fy == sy - 1), ez = (fz == sz - 1);
                const float i0 = (ex) ? 0.f : img[b + 1];
                const float i1 = (ey) ? 0.f : img[b + sx];
                const float i2 = (ex || ey) ? 0.f : img[b + sx + 1];
                const float i3 = (ez) ? 0.f : img[b + sxy];
                const float i4 = (ex || ez) ? 0.f : img[b + sxy + 1];
                const float i5 = (ey || ez) ? 0.f : img[b + sx + sxy];
                const float i6 = (ex || ey || ez) ? 0.f : img[b + sx + sxy + 1];
                const float AA = img[b] + (i0 - img[b]) * vx2;
                const float BB = i1 * vx1 + i2 * vx2;
                const float CC = i3 * vx1 + i4 * vx2;
                const float DD = i5 * vx1 + i6 * vx2;
                *o = (AA * vy1 + BB * vy2) * vz1 + (CC * vy1 + DD * vy2) * vz2;
            }
}

/* tom_peakinc.c:23-52: start at -99999, strict '>', x-fastest scan; pos is
 * 0-based here (the MEX gateway adds 1). */
void orc_find_max(const float *a, const int *dims, float *maxv, int *pos)
{
    long idx = 0;
    *maxv = -99999.0f;
    pos[0] = pos[1] = pos[2] = 0;
    for (int k = 0; k < dims[2]; k++)
        for (int j = 0; j < dims[1]; j++)
            for (int i = 0; i < dims[0]; i++, idx++)
                if (a[idx] > *maxv) {
                    pos[0] = i; pos[1] = j; pos[2] = k;
                    *maxv = a[idx];
                }
}

static inline long wrap(long a, long n) { a %= n; return a < 0 ? a + n : a; }

/*
 * Direct-space FLCF (FP64) of tom_os3_corr for tiny inputs, written from the
 * definitions rather than through FFTs:
 *   mu_I[k]  = sum_u V[k + h - u] M[u] / m          (tom_os3_mean: conv + ifftshift)
 *   E2_I[k]  = sum_u V[k + h - u]^2 M[u] / m        (tom_os3_std)
 *   sig_I[k] = sqrt(max(E2_I - mu_I^2, 0))
 *   c[k]     = sum_u V[k + u - h + s] tm[u] / m     (ifftshift(ifftn(F_I conj F_T)))
 *   score[k] = (c - mu_I mu_T) / (sig_T sig_I)   where sig_T sig_I > 0, else 0
 * with h = floor(n/2) per axis, s = 2 floor(N/2) - N per axis, all indices
 * modulo N (circular FFT wrap), tm = masked template, M = soft mask,
 * m = #{M > 0}. Volumes x-fastest; N = (NX,NY,NZ), n = (nx,ny,nz).
 */
void orc_flcf_direct(const float *V, const long *N, const double *tm, const double *M,
                     const long *n, double m, double mu_T, double sig_T,
                     double *score, double *mu_out, double *sig_out)
{
    const long NX = N[0], NY = N[1], NZ = N[2];
    const long nx = n[0], ny = n[1], nz = n[2];
    const long hx = nx / 2, hy = ny / 2, hz = nz / 2;
    const long sx = 2 * (NX / 2) - NX, sy = 2 * (NY / 2) - NY, sz = 2 * (NZ / 2) - NZ;
    for (long kz = 0; kz < NZ; kz++)
        for (long ky = 0; ky < NY; ky++)
            for (long kx = 0; kx < NX; kx++) {
                double s1 = 0, s2 = 0, c = 0;
                for (long uz = 0; uz < nz; uz++)
                    for (long uy = 0; uy < ny; uy++)
                        for (long ux = 0; ux < nx; ux++) {
                            const long u = ux + nx * (uy + ny * uz);
                            if (M[u] != 0) {
                                const double v = V[wrap(kx + hx - ux, NX) +
                                                   NX * (wrap(ky + hy - uy, NY) +
                                                         NY * wrap(kz + hz - uz, NZ))];
                                s1 += v * M[u];
                                s2 += v * v * M[u];
                            }
                            if (tm[u] != 0) {
                                const double w = V[wrap(kx + ux - hx + sx, NX) +
                                                   NX * (wrap(ky + uy - hy + sy, NY) +
                                                         NY * wrap(kz + uz - hz + sz, NZ))];
                                c += w * tm[u];
                            }
                        }
                const double mu = s1 / m;
                double var = s2 / m - mu * mu;
                if (var < 0) var = 0;
                const double sig = sqrt(var);
                const long o = kx + NX * (ky + NY * kz);
                const double den = sig_T * sig;
                score[o] = den > 0 ? (c / m - mu * mu_T) / den : 0.0;
                if (mu_out) mu_out[o] = mu;
                if (sig_out) sig_out[o] = sig;
            }
}

/* rot2d (Sptrans/tom_rotatec.c:49-142): phi degrees (nphi 1) or a 2x2 matrix
 * rm00 rm10 rm01 rm11 (nphi 4); pivot (px, py) 0-based. */
void orc_rot2d(const float *image, float *rotimg, long sx, long sy, const float *p_phi, int nphi, float px,
               float py)
{
    float rm00, rm01, rm10, rm11;
    if (nphi == 1) {
        float phi = p_phi[0];
        phi = phi * ORC_PI / 180;
        rm00 = cos(phi);
        rm01 = -sin(phi);
        rm10 = sin(phi);
        rm11 = cos(phi);
    } else {
        rm00 = p_phi[0];
        rm10 = p_phi[1];
        rm01 = p_phi[2];
        rm11 = p_phi[3];
    }
    const long sx_1 = sx - 1, sy_1 = sy - 1;
    float pi, pj;
    long i, j;
    for (pj = -py, j = 0; j < sy; j++, pj++) {
        for (pi = -px, i = 0; i < sx; i++, pi++) {
            const float r_x = px + rm00 * pi + rm10 * pj;
            if (r_x < 0 || r_x > sx_1) { *rotimg++ = 0; continue; }
            const float r_y = py + rm01 * pi + rm11 * pj;
            if (r_y < 0 || r_y > sy_1) { *rotimg++ = 0; continue; }
            const long floorx = r_x;
            const float vx2 = r_x - floorx;
            const long floory = r_y;
            const float vy2 = r_y - floory;
            const float *imgp = &image[floorx + sx * floory];
            float AA, BB;
            if (r_x < sx - 1) AA = *imgp + (imgp[1] - *imgp) * vx2;
            else AA = *imgp;
            if (r_y < sy - 1) {
                if (r_x < sx - 1) BB = imgp[sx] + (imgp[sx + 1] - imgp[sx]) * vx2;
                else BB = imgp[sx];
            } else {
                BB = 0;
            }
            *rotimg++ = AA + (BB - AA) * vy2;
        }
    }
}
