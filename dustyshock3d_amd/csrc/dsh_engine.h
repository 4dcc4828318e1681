// dsh_engine.h - host engine: a batch of independent tubes stepped on one GPU stream.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/dsh.h"
#include "dsh_kernels.h"
#include "dsh_types.h"

namespace dsh {

struct HostSpecies {
    int n = 0;
    bool set = false;
    bool device_ic = false;   // generated on the device at prepare() from ic
    TubeIC ic{};
    std::vector<double> x, y, z, vx, vy, vz, m, h, e, rho, p, ts;
    std::vector<int> ind;
};

// Grid rule shared by every entry point (the oracle restates it independently).
void build_grid(const dsh_params& P, double hmin, TubeConst& T);
// Physics, time-step and precision fields of a tube's constants from its parameters.
void tube_physics(const dsh_params& P, TubeConst& C);

class Engine {
public:
    explicit Engine(const std::vector<dsh_params>& params);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int upload(int tube, int species, const dsh_soa_view* v, int pm);
    // Stage a seeded shock-tube species to be generated on the device at prepare().
    int stage_ic(int tube, int species, const TubeIC& ic);
    // Take tube src_tube of a prepared engine, state and clock, without a host round
    // trip: at prepare() its arrays are copied device to device and its grid is kept.
    int adopt(int tube, const Engine& src, int src_tube);
    // Host view of a tube's device arrays in the current buffer (for device-side copies).
    int device_fields(int tube, int species, int* cur_out, long long* off_out) const;
    const PBuf& buffers() const { return a_.B; }
    int prepare();
    // Launch nsteps steps (no host synchronisation in between) and wait for them.
    int run(int nsteps);
    // Launch nsteps steps without waiting (serving loop); sync() waits.
    int launch_steps(int nsteps);
    int sync();
    int download(int tube, int species, const dsh_soa_view* v);
    int neighbour_stats(int tube, int species, int* count, uint64_t* hash);
    // particle indices of one (tube, species) in the device's sorted order of the last step
    int sorted_ids(int tube, int species, int* ids);
    int refresh_dyn();   // copy per-tube dynamic state to the host
    const TubeDyn& dyn(int tube) const { return dyn_h_[tube]; }
    const TubeConst& tube_const(int tube) const { return tc_h_[tube]; }
    int ntubes() const { return (int)params_.size(); }
    bool prepared() const { return prepared_; }
    long long particles(int tube) const;
    size_t device_bytes() const { return dev_bytes_; }

    void set_stats(bool on) { stats_ = on; graph_dirty_ = true; }
    void set_timing(bool on) { timing_ = on; }
    void set_graph(bool on) { use_graph_ = on; }
    void set_pair_variant(int v) {
        pair_variant_ = v;
        a_.variant = v;
        graph_dirty_ = true;
    }
    int pair_variant() const { return pair_variant_; }
    // candidates staged per piece in the tiled pair kernels (kTiles); every size gives the
    // same results bit for bit
    void set_tile(int t) {
        tile_ = t;
        a_.tile = t;
        graph_dirty_ = true;
    }
    int tile() const { return tile_; }
    // waves per block of the tiled pair kernels (kBlockWaves); each wave keeps its own chunk
    // and LDS tile, so every block shape gives the same results bit for bit
    void set_block_waves(int w) {
        waves_ = w;
        a_.waves = w;
        graph_dirty_ = true;
    }
    int block_waves() const { return waves_; }
    // Times steps_per_trial steps of each candidate configuration (block shapes: piece size x
    // waves per block; with variants also the direct pair kernels, whose sums differ in
    // rounding order) on the live state and keeps the fastest.  The trial steps are ordinary
    // steps of the run.
    int autotune(int steps_per_trial, bool with_variants, int* tile, int* variant, double* ms_per_step);
    // per-stage GPU time (ms) averaged over the steps launched with timing on
    const std::vector<std::pair<std::string, double>>& kernel_times() const { return ktimes_; }
    int timed_steps() const { return timed_steps_; }
    const std::string& error() const { return err_; }
    hipStream_t stream() const { return stream_; }
    int error_status();   // DSH_OK or the first tube error as a status

private:
    int fail(int code, const std::string& msg);
    int hip_check(hipError_t e, const char* what);
    template <class T> int alloc(T** p, size_t count);
    int enqueue_step(bool timed);
    int build_graph(int steps, hipGraph_t& graph, hipGraphExec_t& exec);

    std::vector<dsh_params> params_;
    std::vector<HostSpecies> staged_;   // [tube * kMaxSpecies + species]
    struct Migrant {
        const Engine* src = nullptr;
        int tube = -1;
    };
    std::vector<Migrant> migr_;
    std::vector<TubeConst> tc_h_;
    std::vector<TubeDyn> dyn_h_;
    std::vector<Seg> pseg_, gseg_, cseg_, aseg_;
    std::vector<void*> allocs_;
    StepArgs a_{};
    hipStream_t stream_ = nullptr;
    // one step, and kGraphSteps steps, captured as graphs (rebuilt when graph_dirty_)
    static constexpr int kGraphSteps = 8;
    hipGraph_t graph_ = nullptr, graph_n_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr, graph_exec_n_ = nullptr;
    bool prepared_ = false, stats_ = false, timing_ = false, use_graph_ = true, graph_dirty_ = true;
    int precision_ = 0;
    int pair_variant_ = kPairTiled;
    int tile_ = 192;
    int waves_ = 1;
    long long ntot_ = 0;
    size_t dev_bytes_ = 0;
    std::vector<std::pair<std::string, double>> ktimes_;
    std::vector<hipEvent_t> events_;   // (stages + 1) per timed step, pending until sync()
    int pending_timed_ = 0, timed_steps_ = 0;
    int collect_times();
    std::string err_;
};

}  // namespace dsh
