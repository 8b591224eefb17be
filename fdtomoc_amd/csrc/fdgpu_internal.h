// fdgpu_internal.h -- the device context and helpers shared by the host
// translation units of the C ABI (fdgpu.cpp: models, batched solves, tables;
// serve.cpp: the request loop).
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fdtomoc_amd.h"
#include "serve_shared.h"
#include "sweep_solver.h"
#include "table_prep.h"

namespace fdgpu_detail {

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct Model {
    int nx = 0, ny = 0, nz = 0;
    float *d_slow = nullptr;
    size_t bytes = 0;
    double vmin = 0.0, vmax = 0.0;   // velocity range (skip-safety bound)
};

struct Table {
    fdgpu_par par{};
    int model = -1;
    int ox = 0, oy = 0, oz = 0;
    long long n = 0;
    float *d_time = nullptr;
    char *d_aux = nullptr;     // sinq | init_idx | init_dist | init_val | stats
    size_t bytes = 0;
    int *d_stats = nullptr;
    fdsweep::TableJob job{};   // its device descriptor, for re-solves in place (fdgpu_resolve)
    std::vector<int> stats;
    fdprep::Prepared prep;
    bool from_file = false;    // loaded by fdgpu_table_load: no solve inputs behind it
};

// One background table write (fdgpu_tables_write_async): what the file
// needs, captured at submission.
struct WriteJob {
    int table = 0;
    const float *d_time = nullptr;
    long long n = 0;
    fdgpu_par par{};
    fdprep::Prepared prep;
    bool has_head = false;
    unsigned char head[232] = {};
    std::string path;
};

// The background writer: a thread with its own copy stream and two pinned
// buffers, so that table k + 1 is copied while table k is written.
struct Writer {
    std::thread th;
    std::mutex mu;
    std::condition_variable cv, idle;
    std::deque<WriteJob> q;
    int inflight = 0;               // jobs submitted and not yet written
    int err = 0;                    // first error since the last wait
    bool stop = false;
    std::map<int, int> busy;        // table id -> its jobs in flight
    hipStream_t stream = nullptr;
    hipEvent_t done[2] = {nullptr, nullptr};
    float *pinned[2] = {nullptr, nullptr};
    size_t pinned_bytes = 0;
};

// First-fit allocator of byte ranges inside one block (HBM arena or pinned
// staging area); 256-byte granules.
class RangeAlloc {
  public:
    void reset(size_t bytes) {
        free_.clear();
        if (bytes) free_[0] = bytes;
        total_ = bytes;
        used_ = 0;
    }
    // offset of a new range, or SIZE_MAX
    size_t take(size_t bytes) {
        bytes = align256(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_ += bytes;
            return off;
        }
        return SIZE_MAX;
    }
    void give(size_t off, size_t bytes) {
        bytes = align256(bytes ? bytes : 1);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t total() const { return total_; }

  private:
    std::map<size_t, size_t> free_;
    size_t total_ = 0, used_ = 0;
};

// Device geometry tables of one grid (fdsweep::geo_fill_item layout).
struct GeoBuf {
    double *d = nullptr;
    size_t bytes = 0;
};

// One admitted serving request.
struct ServeReq {
    unsigned long long seq = 0;
    fdgpu_par par{};
    long long n = 0;
    size_t a_off = 0, a_bytes = 0;   // HBM arena block
    size_t s_off = 0, s_bytes = 0;   // pinned staging block (freed on completion)
    float *d_time = nullptr;
    int *d_stats = nullptr;
    bool done = false;
    std::vector<int> stats;
    fdprep::Prepared prep;
};

struct Serve {
    bool on = false;
    bool launched = false;             // a server kernel may be running
    hipStream_t sstream = nullptr;     // the server kernel
    hipStream_t cstream = nullptr;     // read-back copies while it runs
    hipEvent_t kdone = nullptr;        // recorded after each server launch
    fdsweep::ServeCtl *ctl = nullptr;  // pinned, coherent
    fdsweep::ServeJob *ring = nullptr; // pinned, coherent, `cap` entries
    unsigned long long *done = nullptr;// pinned, coherent, `cap` entries
    int cap = 0;
    fdsweep::ServeDev *d_dev = nullptr;
    char *arena = nullptr;             // HBM for tables of admitted requests
    RangeAlloc arena_alloc;
    char *stage = nullptr;             // pinned, coherent: request inputs
    RangeAlloc stage_alloc;
    float *bounce = nullptr;           // pinned: read-back buffer
    size_t bounce_bytes = 0;
    char *d_scratch = nullptr;
    fdsweep::Scratch *d_slots = nullptr;
    int nslots = 0;
    long long max_face = 0;
    int maxdim[3] = {0, 0, 0};
    unsigned long long tail = 0;       // requests published
    std::map<int, ServeReq> reqs;
    std::map<unsigned long long, int> pending;   // seq -> request id, not yet seen done
    long long idle_ticks = 0;
    int launches = 0;
};

}  // namespace fdgpu_detail

struct fdgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;   // around one kernel of a sequence (location's node search)
    unsigned long long budget = 0, used = 0;
    int num_cu = 0;
    int wg_per_cu[2] = {1, 1};   // resident workgroups per CU of the wide / narrow sweep kernel
    int lanes = 0;               // workgroup choice (fdgpu_set_workgroup): 0 automatic, 256, 512
    std::map<int, fdgpu_detail::Model> models;
    std::map<int, fdgpu_detail::Table> tables;
    int next_id = 1;
    // per-workgroup scratch of batched solves, grown on demand
    char *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    fdsweep::Scratch *d_slots = nullptr;
    int slots_cap = 0;
    long long slot_face_cap = 0;
    long long slot_node_cap = 0;   // per-slot change stamps: nodes of the largest table
    long long resolved_again = 0;   // tables solved again after a failed consistency check
    long long wave_again = 0;       // ... after a wavefront pass that could not finish
    char *ray_cols = nullptr;       // fdgpu_rays_trace's lane rows (kept zero between calls)
    size_t ray_cols_bytes = 0;
    char *ray_blk = nullptr;        // ... and its other device arrays, kept between calls
    size_t ray_blk_bytes = 0;
    char *ray_out = nullptr;        // ... and its derivatives in output order
    size_t ray_out_bytes = 0;
    bool slot_wave = false;        // the slots hold wavefront state (sized by slot_node_cap)
    int slot_dim_cap = 0;          // per-slot last full-plane faces: largest dimension
    int *d_queue = nullptr;
    fdsweep::TeamCtl *d_teams = nullptr;   // team blocks of batched launches, grown on demand
    int *d_team_first = nullptr;           // ... and the teams' first tickets (after the blocks)
    int teams_cap = 0;
    fdsweep::TableJob *d_jobs = nullptr;   // descriptors of the current batch
    int jobs_cap = 0;
    std::map<long long, hipGraphExec_t> graphs;   // captured batched launches
    // geometry tables of the grids solved so far, keyed by the bits of
    // (ny, nz, h, dq, df, z0r, sin(qc) table); shared by every table of a grid
    std::map<std::vector<unsigned long long>, fdgpu_detail::GeoBuf> geos;
    // scratch of the coarse-to-fine interpolation, grown on demand
    char *d_coarse = nullptr;
    size_t coarse_bytes = 0;
    double last_ms = 0.0;
    // the last fdgpu_locate: passes, pass ms, node-search ms, node-search table bytes
    double loc_timing[4] = {0.0, 0.0, 0.0, 0.0};
    int change_budget = 8;     // per-face schedule knob (fdgpu_set_schedule)
    int wave = 0;              // passes run as wavefronts (fdgpu_set_wave): bit0 first, bit1 reverse
    int team = 0;              // workgroups per large table's wavefront passes (fdgpu_set_team): 0 planned
    int team_share = 0;        // planned: load %; fixed size: tables of at least this % of the mean share
    // this device's copy of the fixed array behind fdgpu_solve_noalloc
    // (sweep_kernel.hip `scratch`), looked up once at fdgpu_open
    char *d_fixed = nullptr;
    fdgpu_detail::Serve serve;
    std::unique_ptr<fdgpu_detail::Writer> writer;   // created by the first background write
    std::mutex mu;
};

namespace fdgpu_detail {

int hip_fail(hipError_t e);
// Background writer: whether `table` has writes in flight; stopping it
// (waits for every write, joins the thread, frees its buffers).
bool writer_busy(fdgpu_ctx *c, int table);
void writer_shutdown(fdgpu_ctx *c);
int dev_alloc(fdgpu_ctx *c, void **p, size_t bytes);
void dev_free(fdgpu_ctx *c, void *p, size_t bytes);
int dev_alloc_resident(fdgpu_ctx *c, void **p, size_t bytes);
void release_ray_columns(fdgpu_ctx *c);
void release_scratch(fdgpu_ctx *c);
// One slot's scratch: face arrays for faces of up to F nodes, plus change
// stamps for tables of up to N nodes and last full-plane marks for dimensions
// up to D (none when N == 0).
size_t slot_bytes(long long F, long long N = 0, int D = 0, bool wave = false);
// Destroys the captured batched launches (their pointers are about to change).
void clear_graphs(fdgpu_ctx *c);
// Carves `ns` scratch slots for faces of up to F nodes out of `base` and
// uploads their descriptors to `d_slots`.
int fill_slots(char *base, int ns, long long F, fdsweep::Scratch *d_slots, long long N = 0, int D = 0,
               bool wave = false);
fdprep::ParValues to_values(const fdgpu_par &p);
// Frees every cached geometry table (stream idle).
void clear_geos(fdgpu_ctx *c);
// A coarse model's validity and fine dimensions (src/c2f.c:262-277), and its
// node coordinates gx | gy | gz in float (c2f.c:258-277).
int coarse_check(const fdgpu_coarse *c, int &nx, int &ny, int &nz);
void coarse_coords(const fdgpu_coarse *c, std::vector<float> &g);
// Host preparation of one table request (validation, geometry, source); `preset`
// receives the srctype 2/3 initial grid.
int prepare_table(const fdgpu_table_desc &d, int model_nx, int model_ny, int model_nz, Table &T,
                  std::vector<float> &preset);
// Serving: stops the request loop and frees its memory (context lock held).
int serve_stop_locked(fdgpu_ctx *c);

}  // namespace fdgpu_detail

// Uploads velocities and converts them to slowness on the device; the caller
// holds the context lock.
int fdgpu_model_update_locked(fdgpu_ctx *c, int model, const float *vel);

