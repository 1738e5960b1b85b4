/*
 * tomhip.h — C ABI of libtomhip.so, the gfx950 (MI355X) implementation of the
 * TOM Toolbox template-matching / subtomogram-scoring hot path.
 *
 * Conventions (those of the reference MEX gateways):
 *   - volumes are float32, x fastest (MATLAB column-major), dims (nx, ny, nz);
 *   - every data pointer may be host or device memory (detected per call);
 *     host pointers are copied in/out around the call (PCIe-inclusive);
 *   - angle indices are 0-based here (MATLAB gateways add 1);
 *   - Euler angles are [phi psi theta] in degrees, ZXZ, tom_rotate order;
 *   - every call returns 0 on success or a negative TOM-style error code
 *     (-1 wrong input, -2 no memory, -3 unsupported, -7 HIP error,
 *     -8 internal) and leaves a message in tomhip_last_error(ctx);
 *     no C++ exception crosses the ABI.
 */
#ifndef TOMHIP_H
#define TOMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tomhip_ctx tomhip_ctx;
typedef struct tomhip_search tomhip_search;

/* search modes */
#define TOMHIP_MODE_FLCF 0 /* tom_os3_findTemplate / tom_os3_corr normalised score */
#define TOMHIP_MODE_XCF 1  /* tom_oscar unnormalised cross-correlation           */

/* ---------------------------------------------------------------- context */
/* One context per device and stream. No reference counterpart (the MEX
 * gateways are stateless); the context owns twiddle tables and workspaces. */
int tomhip_init(int device, tomhip_ctx** ctx);
void tomhip_destroy(tomhip_ctx* ctx);
const char* tomhip_last_error(tomhip_ctx* ctx);
/* Launch on a caller-owned hipStream_t (NULL = the context's own stream). */
int tomhip_set_stream(tomhip_ctx* ctx, void* hip_stream);
int tomhip_synchronize(tomhip_ctx* ctx);

/* ---------------------------------------------------------------- rotation */
/* tom_rotatec(in, out, euler, 'linear', center): replaces the MEX core of
 * Sptrans/tom_rotatec.c:146-401 (gateway :407-480, called from
 * Sptrans/tom_rotate.m:123-128). neuler = 3 ([phi psi theta] deg) or 9
 * (matrix entries rm00 rm10 rm20 rm01 rm11 rm21 rm02 rm12 rm22, i.e. a 3x3
 * MATLAB matrix in column order). center = 0-based pivot (3 floats) or NULL
 * for ceil(n/2). Bit-identical to the reference. */
int tomhip_rotate3d_f32(tomhip_ctx* ctx, const float* src, float* dst, int64_t nx, int64_t ny, int64_t nz,
                        const float* euler, int neuler, const float* center);

/* The 2D branch of tom_rotatec (rot2d, Sptrans/tom_rotatec.c:49-142, gateway
 * :466-475): bilinear rotation of an nx x ny image by phi degrees (nphi = 1)
 * or by a 2x2 matrix rm00 rm10 rm01 rm11 (nphi = 4), pivot center[2]
 * (0-based; tom_rotate passes ceil(size/2)). Pixels whose source falls
 * outside the image are 0; the last row takes no y interpolation, as in the
 * reference. */
int tomhip_rotate2d_f32(tomhip_ctx* ctx, const float* src, float* dst, int64_t nx, int64_t ny, const float* phi,
                        int nphi, const float* center);

/* Batch form: count rotations of one volume, dst[r] at dst + r*nx*ny*nz,
 * euler = [count][3] degrees; optional mask multiplies every output. No
 * reference counterpart (the reference loops tom_rotate per angle,
 * tom_os3_findTemplate.m:264). */
int tomhip_rotate3d_batch_f32(tomhip_ctx* ctx, const float* src, float* dst, int64_t nx, int64_t ny, int64_t nz,
                              const float* euler, int64_t count, const float* center, const float* mask);

/* ---------------------------------------------------------------- peaks */
/* tom_peakinc(in, dims, max, pos): global maximum, first position in x-fastest
 * order, strict '>' from -99999 (Analysis/tom_peakinc.c:23-77, gateway :60).
 * pos is 0-based (the reference MEX adds 1). */
int tomhip_peak_f32(tomhip_ctx* ctx, const float* map, const int32_t dims[3], float* max_val, int32_t pos[3]);

/* Allocation-free form of tomhip_peak_f32. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops such as tom_peakc's per-pick calls, harnesses
 * that forbid hipMalloc). Same result as tomhip_peak_f32; the map must be
 * device memory (anything else returns -1), outputs are caller-owned host
 * scalars. The block partials live in a 16 KiB static device array per GPU
 * (held until the process ends, registered by tomhip_init); calls that
 * overlap on one GPU share it and race. No hipMalloc / hipFree is made. */
int tomhip_peak_f32_noalloc(tomhip_ctx* ctx, const float* map, const int32_t dims[3], float* max_val,
                            int32_t pos[3]);

/* tom_peak(A) (Analysis/tom_peak.m:80-108): the maximum and EVERY position
 * holding it, 0-based, in MATLAB's find order (x fastest); pos [capacity][3];
 * *count = number of ties (-1 error when they exceed capacity). */
int tomhip_peak_all_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], float* max_val, int32_t* pos,
                        int64_t capacity, int64_t* count);
/* tom_peak(A, 'spline'), tom_peak(A, R) and tom_peak(A, R, 'spline')
 * (Analysis/tom_peak.m:117-123, 203-278, 288-430, tom_intpeak :437-525) for
 * 3D volumes: the first maximum (x-fastest order); with spline != 0 each axis
 * is refined by the parabola through the maximum and its two neighbours,
 * sampled at steps of 0.001 (the first best sample; axes at a border keep the
 * integer position and the maximum), val = the largest fitted value; with
 * radius > 0 out receives the map with |d| <= radius around the integer
 * maximum set to 0 (tom_sphere, clipped to the volume). c is 0-based. */
int tomhip_peak_radius_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], int radius, int spline,
                           double c[3], double* val, float* out);

/* ---------------------------------------------------------------- masks, weights */
/* tom_os3_sphereMask(template) (Oscar3/matlabPicker/search/tom_os3_sphereMask.m:4-36):
 * binary r < floor(n/2)-1 for cubic boxes, ones otherwise; *count = #{M>0}. */
int tomhip_spheremask_os3_f32(tomhip_ctx* ctx, const int64_t n[3], float* out, int64_t* count);
/* tom::init_spheremask (src/tom/core/volume_fcn.cpp:1013-1030): 1 inside
 * radius, Gaussian edge exp(-((r-radius)/sigma)^2) up to max_radius (default
 * radius+sqrt(2 sigma) when max_radius < radius); sigma <= 0: binary r < radius. */
int tomhip_spheremask_f32(tomhip_ctx* ctx, const int64_t n[3], float radius, float sigma, float max_radius,
                          float* out);
/* tom_wedge(ones(n), semi_angle) .* tom_bandpass annulus lo <= |k| <= hi on the
 * centred (fftshift) grid (Filtrans/tom_wedge.m:62-77, tom_bandpass.m:43-52).
 * semi_angle_deg <= 0 disables the wedge, hi <= 0 disables the bandpass. */
int tomhip_weight_f32(tomhip_ctx* ctx, const int64_t n[3], double semi_angle_deg, double lo, double hi,
                      float* out);

/* ---------------------------------------------------------------- template prep */
/* tom_norm(1e5*(T+1),'mean0+1std') then tom_os3_modifyImage (box-size W,
 * FP64) and tom_os3_normUnderMask (tom_os3_findTemplate.m:157, 233-234).
 * W: centred weight of the template's size or NULL. */
int tomhip_prepare_template_f32(tomhip_ctx* ctx, const float* tmpl, const int64_t n[3], const float* mask,
                                const float* W, float* tmpl_out, double* mu_t, double* sig_t);

/* ---------------------------------------------------------------- search plan */
/* Exhaustive search state (tom_os3_findTemplate job / tom_oscar loop):
 * F_I and, in FLCF mode, the FP64 local mean/std of the volume under the mask
 * (tom_os3_mean.m:44-63, tom_os3_std.m:43-68) are computed once here.
 * batch = angles per launch group (1..16). */
int tomhip_search_create(tomhip_ctx* ctx, const float* volume, const int64_t N[3], const float* mask,
                         const int64_t n[3], int mode, int batch, tomhip_search** out);
/* tmpl: prepared template (tomhip_prepare_template_f32 output in FLCF mode,
 * raw template in XCF mode); W: centred box weight or NULL; remask = 1 masks
 * the rotated (filtered) template again as tom_os3_corr.m:151 does inside the
 * findTemplate loop, 0 masks it once (tom_os3_corr called directly). Resets
 * the running maxima. */
int tomhip_search_set_template(tomhip_search* s, const float* tmpl, const float* W, double mu_t, double sig_t,
                               int remask);
/* Scan count angles (host [count][3] degrees) whose indices are
 * first_index .. first_index+count-1. May be called repeatedly. */
int tomhip_search_run(tomhip_search* s, const float* euler, int64_t count, int64_t first_index);
/* FLCF: score = (c/m - mu_I mu_T)/(sig_T sig_I) of the best angle (0 where
 * sig_I = 0); XCF: tom_oscar's ccfmax in its orientation. angle: 0-based. */
int tomhip_search_finalize(tomhip_search* s, float* score, int32_t* angle);
/* Per-kernel timing: HIP event pairs recorded around every launch on the
 * search stream (no extra synchronisation); this call resolves them.
 * ms/counts: 8 entries (rotate, box weight, template fwd, pass A, pass B,
 * pass C, -, -). */
int tomhip_search_profile(tomhip_search* s, int enable, double* ms, int64_t* counts);
/* Replay full angle batches through one captured hipGraph (1, default) or
 * launch every kernel eagerly (0). Results are identical. */
int tomhip_search_set_graph(tomhip_search* s, int enable);
/* Column-tile widths (8, 16 or 32; default 16) of the z-FFT pass and the
 * y-FFT pass: workgroup shape only, results are identical. No reference
 * counterpart; set by the autotuner. */
int tomhip_search_set_tiles(tomhip_search* s, int tile_z, int tile_y);
/* Run the y-FFT and c2r/max passes over z-chunks of this many planes (0 =
 * whole volume), so the c2r pass reads the y pass's output back from the
 * memory-side cache. Results are identical. No reference counterpart. */
int tomhip_search_set_zchunk(tomhip_search* s, int planes);
void tomhip_search_destroy(tomhip_search* s);

/* ---------------------------------------------------------------- one-shot drivers */
/* ccc-only tom_os3_findTemplate (Oscar3/matlabPicker/search/tom_os3_findTemplate.m:157-307
 * with tom_os3_bestPeak.m ccc2d): best score and 0-based best angle per voxel.
 * mask: template-sized or NULL (tom_os3_sphereMask); W: centred template-sized
 * weight or NULL; angles [A][3] degrees; batch 0 = default. */
int tomhip_flcf_search_f32(tomhip_ctx* ctx, const float* volume, const int64_t N[3], const float* tmpl,
                           const int64_t n[3], const float* mask, const float* W, const float* angles, int64_t A,
                           int batch, float* score, int32_t* angle);
/* tom_os3_corr(img, template, options) FLCF branch
 * (Oscar3/matlabPicker/search/tom_os3_corr.m:77-192): mask NULL uses
 * tom_os3_sphereMask(template). Includes the low-std branch (:110-125): when
 * the image's mean local std is below 0.01 the template is renormalised as
 * (T*1000+1000).*M, 'mean0+1std', and if the image stays featureless (mean
 * local std below 0.01 of its global std) voxels whose local std is below
 * 0.05 of it score 0. */
int tomhip_os3_corr_f32(tomhip_ctx* ctx, const float* img, const int64_t N[3], const float* tmpl,
                        const int64_t n[3], const float* mask, float* out);
/* tom_oscar(vol, templ, ...) (Filtrans/tom_oscar.m:1-82) over an explicit angle
 * list: ccfmax and the 0-based index of the best angle (-1 never set). */
int tomhip_oscar_f32(tomhip_ctx* ctx, const float* volume, const int64_t N[3], const float* tmpl,
                     const int64_t n[3], const float* angles, int64_t A, float* ccfmax, int32_t* angle);

/* Box-size FLCF by direct sums: tom_os3_corr(search, tmpl) when the search
 * volume has the template's size (Oscar3/matlabPicker/search/tom_os3_corr.m:
 * 77-192, every voxel divided: res = num ./ den), circular; with search =
 * tmpl it is the template autocorrelation that the second-order correlation
 * uses (analyze/tom_os3_analyzePeak.m:180-194; os3_soc, Oscar3/c++Picker/src/
 * tom/os3/os3_correlation.cpp:164-171, whose os3_flcf(pattern, pattern, mask)
 * makes it). mask NULL: tom_os3_sphereMask(template). Edges <= 128. */
int tomhip_flcf_box_f32(tomhip_ctx* ctx, const float* search, const float* tmpl, const float* mask,
                        const int64_t n[3], float* out);
/* tom_os3_bestPeak (Oscar3/matlabPicker/search/tom_os3_bestPeak.m:13-47;
 * the three-score rule is os3_results::update, Oscar3/c++Picker/src/tom/os3/
 * os3_results.cpp:127-150): where every score the rule selects (bit 0 ccc,
 * bit 1 psr, bit 2 soc / autocorrelation) strictly improves on the result
 * maps, copy the orientation's maps and write angle. psr / soc and their
 * result maps may be NULL when the rule does not select them. Result maps
 * start at -1000 (tom_os3_findTemplate.m:215-224). */
int tomhip_os3_bestpeak_f32(tomhip_ctx* ctx, const int64_t N[3], int rule, int32_t angle, const float* ccc,
                            const float* psr, const float* soc, float* res_ccc, float* res_psr, float* res_soc,
                            int32_t* res_angle);

/* ---------------------------------------------------------------- subtomogram scoring */
/* Batched, per-particle local angular search with per-particle wedges, in one
 * persistent launch: the 'peak' mode of tom_mex_corr3d
 * (Oscar3/c++Picker/src/mex/tom_mex_corr3d.cpp:53-806 -> tom::corr3d_batch,
 * src/tom/corr/corr.cpp:969-1400, CorrelationHandlerPeaks
 * src/tom/corr/correlation_handler.cpp:96-120; MATLAB av3_new/tom_av3_corr3d.m)
 * with each particle scanning its own angle list (the local search of
 * av3/alignment/av3_tomomatch.m:173-224).
 *   box_edge[nbox]       box edges, each 32, 48 or 64
 *   templates, masks     packed per box (box k at offset sum_{j<k} B_j^3);
 *                        masks_cc likewise (peak-search masks in unshifted
 *                        correlation coordinates) or NULL for everywhere
 *   wedge_ref_semi_deg[nbox]  WEDGE_REF: tom_wedge semi-angle of each box
 *                        template's missing wedge, turned with the template and
 *                        applied to the particle (corr.cpp:1086-1111); NULL or
 *                        a negative / NaN entry: none
 *   particle_box[P]      box index of each particle; particles packed with
 *                        element offsets particle_offset[P]
 *   wedge_semi_deg[P]    WEDGE_PART: semi-angle of each particle's missing
 *                        wedge (negative / NaN: none; 0: only the cutoff
 *                        sphere, as a scalar 0 does in the gateway);
 *                        SimpleWedge with cutoff (B-1)/2
 *   angles[P][A][3]      [phi psi theta] degrees
 * Outputs: best_angle[P] (0-based, -1 none), peak[P][3] (0-based, unshifted),
 * ccc[P] (NaN when none). */
int tomhip_subtomo_score_batch(tomhip_ctx* ctx, int nbox, const int32_t* box_edge, const float* templates,
                               const float* masks, const float* masks_cc, const float* wedge_ref_semi_deg, int64_t P,
                               const int32_t* particle_box, const float* particles, const int64_t* particle_offset,
                               const float* wedge_semi_deg, const float* angles, int64_t A, int32_t* best_angle,
                               int32_t* peak, float* ccc);

/* ---------------------------------------------------------------- picklist */
/* tom_os3_returnPicklist(peaks, angles, job, n) without binning
 * (Oscar3/matlabPicker/results/tom_os3_returnPicklist.m:40-95): optional
 * tom_os3_eraseBorders, then up to n x tom_peakc(peaks, ceil(tsize[0]/2))
 * (Analysis/tom_peakc.m:39-86) until std(peaks) = 0. angle_map (int32, the
 * search's best-angle map) may be NULL. out: [n][5] rows x y z (1-based)
 * value angle; *count = rows written. The input map is not modified. */
int tomhip_picklist_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], const int32_t* angle_map,
                        const int64_t tsize[3], int64_t n, int erase_borders, float* out, int64_t* count);

/* Local exact refinement of candidates (no reference counterpart as an entry
 * point; it evaluates the same score as tomhip_search_run/finalize, written
 * as a direct sum): for candidate p with group g = cand[p][3], every
 * orientation a of euler[a][g][0..2] (A per group, degrees) is scored at
 * cand[p][0..2] (0-based voxel) and its 26 neighbours. Outputs per candidate:
 * best score, best a (first wins, strict '>'), best voxel as x y z.
 * Needs an FLCF plan with its template set. volume: the searched volume.
 * centroid (optional, 3 floats x y z): the candidates locate the rotated
 * template point p + R d (d in voxels from the rotation centre, e.g. the
 * template's centre of mass) instead of the centre p; orientations are then
 * scored in chunks of 4 around cand - round(mean R d of the chunk). */
int tomhip_search_refine(tomhip_search* h, const float* volume, const float* euler, int G, int A,
                         const int32_t* cand, int64_t P, const float* centroid, float* score, int32_t* angle,
                         int32_t* pos);

/* ---------------------------------------------------------------- secondary scores */
/* Peak-to-sidelobe ratio map of a correlation map (os3_psr,
 * Oscar3/c++Picker/src/tom/os3/os3_correlation.cpp:183-262; MATLAB
 * tom_os3_analyzePeak.m:61-103 'PSR'): psr = (xcf - mean) / std with the local
 * mean and std of xcf under a ring of the template's size n (binary sphere of
 * radius n/2 minus the sphere of radius max(n/4, 5)), FP64 statistics, 0 where
 * the std is 0. The second-order score (os3_soc, :164-171) is tomhip_os3_corr_f32
 * of the xcf map with the template's autocorrelation. */
int tomhip_psr_f32(tomhip_ctx* ctx, const float* xcf, const int64_t N[3], const int64_t n[3], float* out);

/* tom_bin(im, nbin) of a 3D volume (Sptrans/tom_bin.m:38-51): nbin times,
 * each output voxel = the single-precision sum of its 2x2x2 block in
 * tom_bin's term order, divided by 8; odd trailing samples are dropped.
 * out: (N >> nbin) per axis. */
int tomhip_bin_f32(tomhip_ctx* ctx, const float* in, const int64_t N[3], int nbin, float* out);

/* ---------------------------------------------------------------- learned proposal */
/* No counterpart in the reference: a 3-level 3D U-Net (16/32/64 channels, two
 * 3x3x3 convolutions + ReLU per level, 2x2x2 average pool, trilinear x2
 * upsampling, skip concatenation) with a 1x1x1 head giving a particle logit
 * and nbin orientation-bin logits per voxel. BF16 MFMA convolutions with FP32
 * accumulation, FP32 master weights, Adam. It proposes candidate positions and
 * coarse orientations; the exact FLCF then refines them locally.
 * Volumes: float32 x fastest, edges multiples of 4 and >= 8. */
typedef struct tomhip_unet tomhip_unet;
int tomhip_unet_create(tomhip_ctx* ctx, const int64_t dims[3], int nbin, uint64_t seed, int train,
                       tomhip_unet** out);
void tomhip_unet_destroy(tomhip_unet* h);
int tomhip_unet_num_params(tomhip_unet* h, int64_t* n);
/* parameters in layer order (encoder 1a 1b 2a 2b, bottom a b, decoder 2a 2b
 * 1a 1b, head), each layer W[taps][Cout][Cin] then bias[Cout] */
int tomhip_unet_get_params(tomhip_unet* h, float* out);
int tomhip_unet_set_params(tomhip_unet* h, const float* in);
/* Adam state of a training network, for checkpoints: first and second
 * moments m, v in the parameter order above (num_params floats each) and the
 * number of updates taken (bias correction uses it) */
int tomhip_unet_get_optimizer(tomhip_unet* h, float* m, float* v, int64_t* step);
int tomhip_unet_set_optimizer(tomhip_unet* h, const float* m, const float* v, int64_t step);
/* volume and the optional aux map (NULL = zeros) fill input channels 0 and 1 */
int tomhip_unet_forward(tomhip_unet* h, const float* volume, const float* aux);
/* logits [V][80]: channel 0 particle, 1..nbin orientation bins, nbin+1..nbin+3
 * the offset (dx, dy, dz) / 8 from the voxel to the particle centre */
int tomhip_unet_get_logits(tomhip_unet* h, float* out);
/* loss: weighted BCE (positive weight wpos) of channel 0 against tprob[V],
 * averaged over voxels, plus softmax cross-entropy of the bins against
 * tbin[V] (-1 = no label) averaged over labelled voxels, plus (toff optional,
 * [V][4] = dx/8, dy/8, dz/8, weight) the squared error of the offset channels
 * averaged over voxels of weight > 0. Computes all parameter gradients; loss
 * (optional) = {bce, ce, offset}. */
int tomhip_unet_backward(tomhip_unet* h, const float* tprob, const int32_t* tbin, const float* toff, float wpos,
                         double* loss);
int tomhip_unet_get_grads(tomhip_unet* h, float* out);
int tomhip_unet_adam(tomhip_unet* h, float lr);
/* prob[V] = sigmoid(logit 0), bin[V] = argmax of the bin logits (0-based),
 * offsets[V][3] = the offset channels (each output optional) */
int tomhip_unet_predict(tomhip_unet* h, float* prob, int32_t* bin, float* offsets);
/* Whole-volume inference in overlapping tiles of the network's size (halo
 * voxels of context on each side, tile origins on multiples of 4). volume and
 * the optional aux map are each normalised by their global mean and std.
 * flags bit 0: write the particle logit instead of its sigmoid (no
 * saturation at 1, so local maxima stay sharp). offsets (optional) [V][3]. */
int tomhip_unet_infer(tomhip_unet* h, const float* volume, const float* aux, const int64_t N[3], int halo,
                      int flags, float* prob, int32_t* bin, float* offsets);

/* Candidate extraction for the learned proposal (no reference counterpart):
 * local maxima of map (x fastest) above threshold within the cube of
 * half-edge radius (ties to the lowest linear index), as rows x y z (0-based)
 * value, unordered. *count = number found; at most capacity rows written. */
int tomhip_local_maxima_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], int radius, float threshold,
                            int64_t capacity, float* out, int64_t* count);

/* Move each candidate (rows of `stride` int32 starting x y z, 0-based, in
 * place) to the largest value of map within the cube of half-edge radius
 * around it (clipped to the volume, ties to the lowest index). No reference
 * counterpart: it recentres proposals on a matched-filter map. */
int tomhip_snap_maxima_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], int radius, int32_t* cand,
                           int64_t P, int stride);

/* ---------------------------------------------------------------- EM files */
/* EM format header (Oscar3/c++Picker/include/tom/core/io.h:49-63). */
typedef struct tomhip_em_header {
  int8_t machine; /* 6 = PC little-endian; 0, 3, 5 big-endian */
  int8_t byte2;
  int8_t byte3;
  int8_t type; /* 1 int8, 2 int16, 4 int32, 5 float, 8 complex, 9 double, 10 complex double */
  uint32_t dims[3];
  int8_t comment[80];
  int32_t emdata[40];
  int8_t userdata[256];
} tomhip_em_header;
/* tom_emread (IOfun/tom_emread.m:153-220): header, and the data of any real
 * type converted to float32 (x fastest; capacity in elements). Host code; no
 * context needed. */
int tomhip_em_read_header(const char* path, tomhip_em_header* hdr);
int tomhip_em_read_f32(const char* path, float* out, int64_t capacity, tomhip_em_header* hdr_out);
/* tom_emwrite: float32 PC-format file; hdr (optional) supplies comment,
 * emdata and userdata. */
int tomhip_em_write_f32(const char* path, const float* data, const int64_t dims[3], const tomhip_em_header* hdr);
const char* tomhip_em_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TOMHIP_H */
