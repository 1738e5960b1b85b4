// Exhaustive FLCF template-matching engine (host side of search.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "common.h"
#include "fft_kernels.h"

namespace tomhip {

enum SearchMode : int {
  MODE_FLCF = 0,  // tom_os3_findTemplate / tom_os3_corr normalised score
  MODE_XCF = 1,   // tom_oscar unnormalised cross-correlation
};

struct KernelTimes {
  // accumulated milliseconds and launch counts of the per-angle kernels
  double ms[8] = {0};
  long n[8] = {0};
};
enum KernelId : int { K_ROT = 0, K_BOX = 1, K_TFWD = 2, K_PASSA = 3, K_PASSB = 4, K_PASSC = 5 };

// FP64 local mean / std of a volume under a template-sized mask (device
// pointers; mu, sig: X*Y*Z doubles), as tom_os3_mean / tom_os3_std with
// m = #{mask > 0}; variance at or below 1e-12 x the global variance -> std 0.
void local_stats(Twiddles& tw, hipStream_t s, const float* volume, int X, int Y, int Z, const float* mask, int nx,
                 int ny, int nz, double* mu, double* sig);

class Search {
 public:
  // volume: device, x-fastest X*Y*Z floats. mask: device nx*ny*nz (FLCF mode;
  // may be null in XCF mode). mask_count = #{mask > 0}.
  Search(Twiddles& tw, hipStream_t s, const float* volume, int X, int Y, int Z, const float* mask, int nx,
         int ny, int nz, int mode, int batch);
  ~Search();
  Search(const Search&) = delete;
  Search& operator=(const Search&) = delete;

  // Prepared template (device, nx*ny*nz, already tom_norm'ed in FLCF mode),
  // optional ifftshift-ed box weight (device, nx*ny*nz or null), fixed
  // template statistics.
  // remask: mask the rotated (filtered) template a second time, as
  // tom_os3_corr does inside the findTemplate loop.
  void set_template(const float* tmpl, const float* wsh, double mu_t, double sig_t, bool remask);
  // Rotation matrices (host, [count][9]) for angles first_index..+count-1.
  void run(const float* rm_host, long count, long first_index);
  // Writes score (float) and 0-based best-angle index maps (device or host).
  void finalize(float* score, int32_t* angle);

  // Exact FLCF of a local orientation set around candidates (FLCF mode, after
  // set_template): rm_host [A][G][9] holds A orientations for each of G
  // groups; cand_dev (device) [P] of {x, y, z, group}; each candidate is
  // scored at its position and 26 neighbours for the A orientations of its
  // group. Outputs (device, [P]): best score, best orientation index (0..A-1,
  // first wins), best voxel (linear index). volume: the searched volume.
  // centroid (host, 3 floats, optional): the candidates locate R d + p
  // (the rotated template point d, in voxels from the rotation centre, of a
  // particle at p) rather than p; each chunk of orientations is then scored
  // around candidate - round(mean R d) of the chunk.
  void refine(const float* volume, const float* rm_host, int G, int A, const int32_t* cand_dev, long P,
              const float* centroid, float* bscore, int32_t* bangle, int32_t* bpos);

  // Column-tile widths of pass A (z FFTs) and pass B (y FFTs): 8, 16 or 32
  // columns per workgroup (default 16). Results do not depend on them.
  void set_tiles(int tile_a, int tile_b);
  // Run passes B and C over z-chunks of this many planes (0 = whole volume)
  // so that C reads B's output from the memory-side cache. Results are
  // identical.
  void set_zchunk(int planes);
  bool tiles_ok(int tile_a, int tile_b) const;

  // Queue the plan's later work on stream s, after what it has queued so far.
  void bind_stream(hipStream_t s);

  void set_profile(bool on) { profile_ = on; }
  // FP64 local standard deviation of the volume under the mask (device, N).
  const double* local_std() const { return sig_; }
  // Replay full angle batches through one captured hipGraph (default on;
  // profiling runs eager so every launch can be timed).
  void set_graph(bool on) { use_graph_ = on; }
  // Resolves the recorded event pairs (synchronises the stream) and returns
  // the accumulated per-kernel times.
  const KernelTimes& times();
  void reset_times() { times_ = KernelTimes(); }
  int batch() const { return KA_; }

  // sizes
  int X, Y, Z, Xh, nx, ny, nz, mode;
  long N;
  double m = 1.0, mu_t = 0.0, sig_t = 1.0;
  double mask_sum_ = 0.0;  // sum of the mask values

 private:
  void precompute_stats(const float* volume);
  void release();
  void run_batch(const float* rm_dev, int ka, long first_index, const long* desc);
  void rotate_templates(const float* rm_dev, int count, float* tout, const long* desc, long rm_step);
  void box_filter(int count, float* t, c32* work);
  void launch_passA(int ka);
  void launch_passB(int ka, int z0, int zn);
  void launch_passC(int ka, int z0, int zn, long first_index, const long* desc);
  void capture_graph();
  void drop_graph();
  template <class F>
  void timed(int kid, F&& f);

  Twiddles& tw_;
  hipStream_t s_;
  hipEvent_t swap_ev_ = nullptr;
  int KA_;
  bool profile_ = false;
  KernelTimes times_;
  struct EvRec {
    int kid;
    hipEvent_t a, b;
  };
  std::vector<EvRec> evpool_;  // recorded around launches, resolved lazily
  size_t evused_ = 0;

  // device buffers
  const float* mask_ = nullptr;
  c32* FI_ = nullptr;     // [Z][Y][Xh]   F_I, kx < X/2
  c32* FIn_ = nullptr;    // [Y][Z]       F_I at kx = X/2
  double* mu_ = nullptr;  // [N] local mean (FLCF)
  double* sig_ = nullptr; // [N] local std  (FLCF)
  float* maxc_ = nullptr; // [N] running max of the raw correlation
  int32_t* argm_ = nullptr;  // [N] best angle index
  float* tmpl_ = nullptr;    // [n3] prepared template (owned copy)
  float2* tpairs_ = nullptr; // [n3] its x-neighbour pairs for the rotation
  float* wsh_ = nullptr;     // [n3] ifftshift(W) or null
  float* rm_ = nullptr;      // [cap][9] rotation matrices of the current run
  long rm_cap_ = 0;
  float* t1_ = nullptr;      // [KA][n3] rotated (masked, filtered) templates
  c32* tbox_ = nullptr;      // [KA][n3] box filter workspace
  // refinement workspaces (templates, box filter), kept between refine calls
  // and grown on demand; freed with the plan
  float* ref_tt_ = nullptr;
  c32* ref_work_ = nullptr;
  long ref_cap_ = 0;         // templates of n3 voxels each buffer holds
  c32* Ty_ = nullptr;        // [KA][nz][Y][Xh]
  c32* Tyn_ = nullptr;       // [KA][nz][Y]
  c32* Tx_ = nullptr;        // [KA][nz][ny][pitch] template x-DFT rows, when they do not fit in Q
  c32* Q_ = nullptr;         // [KA][Z][Y][Xh]
  bool started_ = false;
  bool remask_ = true;
  // graph of one full batch, replayed with the device descriptor desc_
  bool use_graph_ = true;
  int tile_a_ = 16, tile_b_ = 16;
  int zchunk_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  long* desc_ = nullptr;  // {batch, first index}
  bool graph_has_w_ = false, graph_remask_ = true;
};

}  // namespace tomhip
