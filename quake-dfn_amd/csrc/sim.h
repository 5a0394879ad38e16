// Host side of a device-resident run (qdfn_sim_t), shared by sim.hip and the
// batched ensemble (ensemble.hip).
#pragma once
#include "dense.h"
#include "hmat.h"
#include "sim_device.h"

// A captured sequence of attempts and its executable graph.
struct GraphSlot {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int attempts = 0;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        attempts = 0;
    }
};

// One run: its operator (borrowed), parameters, stream, device arrays and
// state block.  A run taken into an ensemble is advanced by the ensemble's
// launches; its getters keep working.
struct qdfn_sim {
    qdfn_dense* op = nullptr;   // dense coupling (main) ...
    qdfn_hmat* hop = nullptr;   // ... or H-matrix coupling (main_H)
    qdfn_sim_params P{};
    hipStream_t stream = nullptr;
    qdfn::DevBuf consts, state2, single, ints, cand, ext, hist, histsw, trace, dstate, ranlog, parts;
    qdfn::SimArrays A{};
    qdfn::SimState* d_st = nullptr;
    qdfn::SimState h_st{};
    bool started = false;
    // captured attempt sequences of 1, 2, 4, ... attempts and of graph_k (the
    // run's batch size), so that a run near its step target replays no more
    // attempts than it can use
    std::vector<GraphSlot> graphs;
    int graph_k = 0;
    // timing of coupling kernels (mode 1)
    double coup_ms_sum = 0.0;
    int64_t coup_count = 0;
    int64_t rec_flushed = 0;  // records already handed to a record callback
    // checkpoints of qdfn_sim_run_checkpointed: the run state (V, D, TH, F,
    // SIG both halves, Accel: 11 n doubles) and the state block before steps
    // 0, every, 2 every, ...
    qdfn::DevBuf ckpt;
    std::vector<qdfn::SimState> ck_state;
    int64_t ck_every = 0;
};

namespace qdfn {
// A run on the same operator with the same constants, external tables and
// parameters as s, recording every step into record_capacity records (no
// trace): the replay run of checkpointed recomputation.
qdfn_sim* sim_clone(const qdfn_sim* s, int64_t record_capacity);
// Sets run r's state to checkpoint c of s (r a clone of s).
int sim_restore(const qdfn_sim* s, int64_t c, qdfn_sim* r);
}  // namespace qdfn

