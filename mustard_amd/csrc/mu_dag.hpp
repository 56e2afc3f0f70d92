// mu_dag.hpp -- host-side task recording and lowering to device units.
//
// Recording follows the reference's TiledGraphCreator (mustard/mustard.h:
// 132-384): a task names the tile it writes and the tiles it reads; its
// dependencies are the sorted, de-duplicated last writers of those tiles
// (getSubgraphDependencies, mustard/mustard.h:279-296).  Between begin and end
// of a task the caller issues tile operations (our replacements of the cuBLAS
// / cuSOLVER calls); they are recorded as OpRecs instead of being captured
// from a stream.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "mu_types.h"

namespace mu {

// One tile operation at task level (before lowering).
struct OpRec {
    int32_t op = OP_NONE;
    int32_t flags = 0;
    int32_t m = 0, n = 0, k = 0;  // GEMM m,n,k; SYRK n (in m and n), k; TRSM: m = order, n = nrhs
    int32_t lda = 0, ldb = 0, ldc = 0;
    double alpha = -1.0, beta = 1.0;
    const double *a = nullptr;
    const double *b = nullptr;
    double *c = nullptr;
    int32_t *info = nullptr;
    int32_t info_base = 0;
    // a trailing update at the outer level of a two-level nested
    // factorization (runs on the bulk lane like a recorded GEMM/SYRK task)
    bool update = false;
};

typedef std::pair<int, int> Tile;

struct TaskRec {
    Tile w;
    std::vector<Tile> r;
    std::vector<int> deps;  // sorted unique task ids
    std::vector<OpRec> ops;
    int tag = 0;            // problem id (batched runs)
    // Nested sub-DAG run after `ops` (insertSubgraph): tasks numbered within
    // `sub`, their deps refer to earlier entries of `sub`.
    std::vector<TaskRec> sub;
};

class Dag {
  public:
    // beginCaptureOperation / endCaptureOperation (mustard/mustard.h:148-181).
    int begin_task(Tile w, const std::vector<Tile> &r);
    int end_task();
    int add_op(const OpRec &op);
    void add_edge(int src, int dst);  // insertDependencyKernel (mustard.h:197-211)
    // insertSubgraph (mustard.h:214-248): the tasks of `child` become a
    // nested sub-DAG of the last ended task, run after its own operations.
    int insert_sub(const Dag &child);
    int current() const { return cur_; }
    void set_tag(int tag) { tag_ = tag; }

    std::vector<TaskRec> tasks;
    std::string err;

  private:
    std::map<std::tuple<int, int, int>, int> last_writer_;  // (tag, row, col) -> task
    int cur_ = -1;
    int tag_ = 0;
};

struct LowerOpts {
    int small_max = SMALL_MAX;  // diagonal factorizations above this nest into sub-tiles
    int trsm_max = TRSM_NEST;   // triangular solves above this nest into blocks (<= TRSM_NMAX)
    int trsm_inv_min = SMALL_MAX + 1;  // triangles of at least this order use inverted diagonal blocks
    int nprio = NPRIO;
    int gemm_group = 1;  // GEMM items of gemm_group x gemm_group output blocks (the persistent
                         // kernel pays its per-item scheduling and ordering once per group)
    // Factorizations and solves of order above 2 * nest_outer first split
    // into blocks of about nest_outer (whose diagonal operations nest again),
    // so a whole-matrix task lowers like the tiled DAG instead of into
    // (n / small_max)^2 sub-tiles.
    int nest_outer = 1000;
};

struct Lowered {
    std::vector<UnitDesc> units;
    std::vector<std::vector<uint32_t>> succ;  // per unit
    std::vector<int> task_first_unit;         // per task: first unit id
    std::vector<int> task_nunits;
    uint64_t nitems = 0;
    double flops = 0.0;  // algorithmic flop count of the recorded ops
    double critical_path = 0.0;
    // Workspace of the inverted diagonal blocks and of the F_A32 GEMMs split
    // over k-slices: slot s holds ws_doubles[s] doubles; the executor
    // allocates them, zeroed, and points the TRINV unit (c) and its TRSM unit
    // (b), or the GEMM unit (part), of each slot there.
    std::vector<size_t> ws_doubles;
    std::vector<char> update_unit;  // per unit: lowered from an OpRec with `update`
};

// Point the units of every workspace slot into ws (ws_total_doubles(low)
// doubles) and return the offsets.
size_t ws_total_doubles(const Lowered &low);
void patch_workspace(std::vector<UnitDesc> &units, const Lowered &low, double *ws);

// Lower the recorded DAG to device units (nested sub-DAGs for large
// factorizations and solves), assign priorities and number the items.
bool lower(const Dag &dag, const LowerOpts &opts, Lowered &out, std::string &err);

// Number of items of a unit as the device decodes it.
int unit_items(const UnitDesc &u);

// Algorithmic flops of one recorded operation.
double op_flops(const OpRec &op);

}  // namespace mu
