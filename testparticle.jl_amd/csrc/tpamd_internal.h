// tpamd_internal.h -- declarations shared by the library's translation units:
// the opaque handle types of the C ABI, the error helper, the chunk record of
// batched launches, and the launch entry points each kernel family's unit
// defines (one unit per family, so they compile in parallel).
#pragma once

#include "tpamd_device.h"
#include "tpamd_adjoint.h"
#include "tpamd.h"

#include <hip/hip_runtime.h>

#include <string>

struct tpamd_field {
    tpamd::DevField d;
    void* dev_data;          // owned device buffer (grids)
    void* cells_data;        // owned cell-replicated copy (grids, optional)
    const tpamd_field* src;  // UxB: the B field it was built from
    size_t bytes;
    tpamd::MeshGrid mesh_host;      // Mesh: host copy of the descriptor (device pointers)
};

struct tpamd_ctx {
    int device;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    std::string err;
    // staging buffers for tpamd_solve (grown on demand)
    void* d_buf;
    size_t d_buf_bytes;
    unsigned long long* d_steps;
    // the device address of this device's copy of the allocation-free
    // entry point's scratch array (looked up at tpamd_init; null if that failed)
    unsigned char* scratch;
};

namespace tpamd {

// tpamd_sort.hip: radix sort of (key, record) pairs on key bits [begin, end)
template <typename R>
hipError_t adj_sort_records(void* d_temp, size_t& temp_bytes, const unsigned* keys_in,
                            unsigned* keys_out, const AdjRec<R>* vals_in, AdjRec<R>* vals_out,
                            long long n, int begin_bit, int end_bit, hipStream_t stream);

// tpamd_sort.hip: the bin sums over block-sorted records
template <typename R>
hipError_t adj_bins_seg(const unsigned* keys, const AdjRec<R>* vals, const unsigned* seg_start,
                        const unsigned* seg_count, int nblk, unsigned nbins, unsigned ncell,
                        double* G, hipStream_t stream);
// tpamd_sort.hip: block-local sort of the records by bin (see AdjPart)
template <typename R>
hipError_t adj_localsort_records(const unsigned* keys_in, const AdjRec<R>* vals_in, long long M,
                                 unsigned nbins, unsigned* keys_out, AdjRec<R>* vals_out,
                                 unsigned* seg_start, unsigned* seg_count, hipStream_t stream);

// records the message on ctx (or as the last tpamd_init error) and returns code
int fail(tpamd_ctx* ctx, int code, const char* fmt, ...);

// ---------------------------------------------------------------------------
// Batched ensembles: a persistent grid of waves pulls chunks of particles from
// an atomic queue.  A chunk never straddles a field instance, so the
// instance's parameter block is wave-uniform; within a chunk the 64 lanes
// stride over its particles (particles per thread = chunk size / 64), which
// evens out uneven lifetimes.  Every wave leaves once the queue is drained.
// ---------------------------------------------------------------------------
// The batched kernel's chunk list is dealt into kQueues interleaved slices
// (slice q: chunks q, q + kQueues, ...), each with its own claim counter on
// its own 128-B line.  A wave starts on the slice of its block's XCD
// (blocks are dispatched to the 8 XCDs round-robin) and moves on to the
// others when it is drained, so all waves finish; one shared counter
// serialised ~8 claims per microsecond.
constexpr int kQueues = 8;
constexpr int kQueueStride = 32;   // unsigned words between counters (128 B)

struct Chunk {
    int inst;
    int pad;
    long long begin, end;
};

// one launch of a solve kernel: a thread per particle (batch == false, over
// *p) or the persistent chunk-queue kernel over a parameter table (with
// `occ` set, only the kernel's resident blocks per CU are reported)
struct LaunchReq {
    bool batch;
    int adaptive;
    dim3 g, b;
    hipStream_t st;
    const KParams* p;
    const KParams* table;
    const Chunk* chunks;
    int nchunks;
    unsigned* counter;
    unsigned long long* steps;
    int* occ;
    bool rec;                 // records carry E, B or work columns
};

// generic kernels, FP64 and FP32C, per field-kind class: analytic fields
// (tpamd_gen64a/32a.hip), + grids (tpamd_gen64/32.hip), + meshes
// (tpamd_gen64m/32m.hip)
void launch_generic64_analytic(const LaunchReq& r, int alg, bool special);
void launch_generic32_analytic(const LaunchReq& r, int alg, bool special);
void launch_generic64(const LaunchReq& r, int alg);
void launch_generic32(const LaunchReq& r, int alg);
void launch_generic64_mesh(const LaunchReq& r, int alg);
void launch_generic32_mesh(const LaunchReq& r, int alg);
// specialised gridded-B kernels (tpamd_spec64/32.hip): emode kEUniform /
// kEUxB, ns the subcycle class, per a periodic grid, f32 float storage,
// cells the cell-replicated layout
void launch_spec64(const LaunchReq& r, int emode, int ns, bool per, bool f32, bool cells);
void launch_spec32(const LaunchReq& r, int emode, int ns, bool per, bool cells);
// the reference-form multistep classes (ns = kRefBase + n)
void launch_spec64r(const LaunchReq& r, int emode, int ns, bool per, bool f32, bool cells);
void launch_spec32r(const LaunchReq& r, int emode, int ns, bool per, bool cells);
// full-orbit segments of the hybrid solver (tpamd_hybrid.hip)
struct FoArgs {
    double* state;       // [N][8] x, v, t, dt
    long long* iters;    // [N][2] it, steps
    int* flags;          // [N] in: TPAMD_FO_ENTER | TPAMD_FO_CHECKED; out: status
    double* sync;        // [N][6]
    double dtmin, dtmax, safety;
    long long check_interval, maxiters;
};
void launch_fo_segment(const KParams& p, const FoArgs& a, int fk, hipStream_t st);
// the checkpointed adjoint (tpamd_adj.hip)
int adjoint_launch(tpamd_ctx* ctx, const KParams& p, const tpamd_field* B, const double* xv0,
                   int64_t N, const double* targets, int K, int64_t nt, double* loss,
                   double* dL_dB, double* dL_dxv0, tpamd_stats* stats, bool f32, bool bf32,
                   bool per, bool cells);

}  // namespace tpamd

#define HIP_TRY(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return tpamd::fail((ctx), e_ == hipErrorOutOfMemory ? TPAMD_ERR_OOM : TPAMD_ERR_DEVICE, \
                               "%s failed: %s", #expr, hipGetErrorString(e_));               \
    } while (0)
