// lu_mustard -- tiled LU without pivoting on one MI355X.
//
// Command line, task order and output lines of mustard/lu_mustard.cu
// (tiledLU 256-670, main 682-716); the tile operations are the library's
// gfx950 kernels instead of cuSOLVER / cuBLAS, and --subgraph runs the
// persistent device scheduler instead of device-launched CUDA graphs.
#include "cli.hpp"

int myPE = 0;

static cli::MustardConfig cfg;

// lu_mustard.cu:133-254: the tiled LU of one diagonal tile (order B, leading
// dimension N) in subT x subT sub-tiles, recorded as its own graph for
// insertSubgraph.  The creator is built on a fresh graph handle and ends
// without being instantiated, so its tasks stay under that handle.
static hipGraph_t recordSubgraph(double *subMatrix, int subT, hipStream_t s, int *d_info)
{
    const int N = (int)cfg.N, subN = (int)cfg.B, subB = subN / subT;
    std::cout << "PE " << myPE << ". Create subgraph with subN=" << subT << " (" << subT
              << " of " << subB << "x" << subB << " tiles)" << std::endl;
    auto getMatrixBlock = [&](double *matrix, int i, int j) {
        return matrix + (size_t)i * subB + (size_t)j * subB * N;
    };
    int totalNodes = subT;
    for (int k = 0; k < subT; k++)
        for (int i = k + 1; i < subT; i++)
            totalNodes += 2 + (subT - (k + 1));
    hipGraph_t subGraph;
    checkHip(hipGraphCreate(&subGraph, 0));
    mustard::TiledGraphCreator creator(s, subGraph, false, totalNodes);
    for (int k = 0; k < subT; k++) {
        creator.beginCaptureOperation(std::make_pair(k, k), {std::make_pair(k, k)});
        checkMustard(mu_dgetrf_nopiv(s, subB, getMatrixBlock(subMatrix, k, k), N, d_info));
        creator.endCaptureOperation();
        for (int i = k + 1; i < subT; i++) {
            creator.beginCaptureOperation(std::make_pair(k, i),
                                          {std::make_pair(k, k), std::make_pair(k, i)});
            checkMustard(mu_dtrsm(s, 'L', 'L', 'N', 'U', subB, subB, 1.0,
                                  getMatrixBlock(subMatrix, k, k), N,
                                  getMatrixBlock(subMatrix, k, i), N));
            creator.endCaptureOperation();
        }
        for (int i = k + 1; i < subT; i++) {
            creator.beginCaptureOperation(std::make_pair(i, k),
                                          {std::make_pair(k, k), std::make_pair(i, k)});
            checkMustard(mu_dtrsm(s, 'R', 'U', 'N', 'N', subB, subB, 1.0,
                                  getMatrixBlock(subMatrix, k, k), N,
                                  getMatrixBlock(subMatrix, i, k), N));
            creator.endCaptureOperation();
            for (int j = k + 1; j < subT; j++) {
                creator.beginCaptureOperation(
                    std::make_pair(i, j),
                    {std::make_pair(i, k), std::make_pair(k, j), std::make_pair(i, j)});
                checkMustard(mu_dgemm(s, 'N', 'N', subB, subB, subB, -1.0,
                                      getMatrixBlock(subMatrix, i, k), N,
                                      getMatrixBlock(subMatrix, k, j), N, 1.0,
                                      getMatrixBlock(subMatrix, i, j), N));
                creator.endCaptureOperation();
            }
        }
    }
    return subGraph;
}

static void tiledLU(const cli::Args &cmdl, bool subgraph)
{
    const size_t N = cfg.N, B = cfg.B, T = cfg.T;
    void *sv = nullptr;
    checkMustard(mu_stream_create(&sv));
    hipStream_t s = (hipStream_t)sv;

    double *d_matrix = nullptr, *d_work = nullptr;
    int *d_info = nullptr;
    checkMustard(mu_malloc((void **)&d_matrix, N * N * sizeof(double)));
    checkMustard(mu_malloc((void **)&d_info, sizeof(int)));
    std::vector<double> h;
    if (cmdl.param({"gen"}) && *cmdl.param({"gen"}) == "spd")
        checkMustard(mu_malloc((void **)&d_work, N * N * sizeof(double)));
    cli::makeMatrix(cmdl, N, h, d_matrix, d_work, s);
    if (d_work)
        checkMustard(mu_free(d_work));

    auto getMatrixBlock = [&](double *matrix, size_t i, size_t j) {
        return matrix + i * B + j * B * N;  // lu_mustard.cu:278-281
    };

    int totalNodes = (int)T;  // lu_mustard.cu:322-326
    for (size_t k = 0; k < T; k++)
        for (size_t i = k + 1; i < T; i++)
            totalNodes += 2 + (int)(T - (k + 1));
    if (cfg.verbose)
        std::cout << "totalNodes=" << totalNodes << std::endl;

    mustard::TiledGraphCreator creator(s, nullptr, subgraph, totalNodes);
    const int n = (int)N, b = (int)B;
    for (size_t k = 0; k < T; k++) {
        // A[k][k] = GETRF(A[k][k])
        creator.beginCaptureOperation(std::make_pair((int)k, (int)k), {std::make_pair((int)k, (int)k)});
        // lu_mustard.cu:358-388: with --subgraph a tile above MAX_TILE is
        // factored by a nested sub-graph of ceil(B / MAX_TILE)^2 sub-tiles
        // (when they divide the tile evenly; otherwise the library's own
        // nesting of the whole-tile GETRF does the same job)
        const int subT = (int)((B + MAX_TILE - 1) / MAX_TILE);
        const bool nested = subgraph && B > MAX_TILE && B % subT == 0;
        if (!nested)
            checkMustard(mu_dgetrf_nopiv(s, b, getMatrixBlock(d_matrix, k, k), n, d_info));
        creator.endCaptureOperation();
        if (nested) {
            hipGraph_t subLU = recordSubgraph(getMatrixBlock(d_matrix, k, k), subT, s, d_info);
            creator.insertSubgraph(subLU);
            mustard::releaseGraph(subLU);
            checkHip(hipGraphDestroy(subLU));
        }
        for (size_t i = k + 1; i < T; i++) {
            // U[k][i] = L[k][k]^-1 A[k][i]
            creator.beginCaptureOperation(std::make_pair((int)k, (int)i),
                                          {std::make_pair((int)k, (int)k), std::make_pair((int)k, (int)i)});
            checkMustard(mu_dtrsm(s, 'L', 'L', 'N', 'U', b, b, 1.0, getMatrixBlock(d_matrix, k, k),
                                  n, getMatrixBlock(d_matrix, k, i), n));
            creator.endCaptureOperation();
        }
        for (size_t i = k + 1; i < T; i++) {
            // L[i][k] = A[i][k] U[k][k]^-1
            creator.beginCaptureOperation(std::make_pair((int)i, (int)k),
                                          {std::make_pair((int)k, (int)k), std::make_pair((int)i, (int)k)});
            checkMustard(mu_dtrsm(s, 'R', 'U', 'N', 'N', b, b, 1.0, getMatrixBlock(d_matrix, k, k),
                                  n, getMatrixBlock(d_matrix, i, k), n));
            creator.endCaptureOperation();
            for (size_t j = k + 1; j < T; j++) {
                // A[i][j] -= L[i][k] U[k][j]
                creator.beginCaptureOperation(
                    std::make_pair((int)i, (int)j),
                    {std::make_pair((int)i, (int)k), std::make_pair((int)k, (int)j), std::make_pair((int)i, (int)j)});
                checkMustard(mu_dgemm(s, 'N', 'N', b, b, b, -1.0, getMatrixBlock(d_matrix, i, k), n,
                                      getMatrixBlock(d_matrix, k, j), n, 1.0,
                                      getMatrixBlock(d_matrix, i, j), n));
                creator.endCaptureOperation();
            }
        }
    }
    if (subgraph && cfg.verbose)
        creator.printDeps();

    const bool stream_mode = cmdl.flag({"stream"});
    mu_exec_opts_t opts;
    std::memset(&opts, 0, sizeof opts);
    cmdl.get({"grid"}, opts.grid);
    const std::string *trace = cmdl.param({"trace"});
    opts.trace = trace != nullptr;
    // --sm: the CU budget of each task in the persistent scheduler
    // (FLAGS_OCCUP, mustard.h:82-85); the tiled graph's level kernels mix the
    // items of many tasks and run without one
    opts.sm = subgraph ? cfg.smLimit : 0;
    mustard::Executable ex = creator.instantiate(
        stream_mode ? mustard::Mode::Stream
                    : (subgraph ? mustard::Mode::Subgraph : mustard::Mode::Tiled),
        &opts);
    if (cmdl.flag({"dot"}))
        creator.writeDot("./graph.dot");
    if (cfg.verbose) {
        mu_exec_stats_t st = ex.stats();
        printf("lowered: %lld units, %lld items, %lld edges, grid %d\n", st.units, st.items,
               st.edges, st.grid);
    }
    std::vector<double> times = cli::timedRuns(cmdl, cfg, ex, d_matrix, h, d_info, s);
    if (trace && subgraph)
        cli::writeTrace(*trace, ex);
    const int rc = cli::finish(cmdl, cfg, false, times, h, d_matrix, d_info, s, ex.stats().flops);
    checkMustard(mu_free(d_matrix));
    checkMustard(mu_free(d_info));
    checkMustard(mu_stream_destroy(sv));
    if (rc)
        std::exit(rc);
}

int main(int argc, char **argv)
{
    cli::Args cmdl = cli::parse(argc, argv);
    if (!cli::parseCommonArgs(cmdl, cfg)) {
        cli::printSingleNodeUsage(argv[0], "LU");
        return 1;
    }
    cli::initDevice(cmdl, cfg);
    myPE = cfg.myPE;
    if (!(cmdl.flag({"tiled"}) || cmdl.flag({"subgraph"})))
        cfg.T = 1;
    cfg.B = cfg.N / cfg.T;
    cli::banner(cmdl, cfg);
    // Single-kernel mode (no --tiled / --subgraph) is one task on the whole
    // matrix, nested by the library and run as a level-batched hipGraph.
    const bool subgraph = cmdl.flag({"subgraph"});
    tiledLU(cmdl, subgraph);
    return 0;
}
