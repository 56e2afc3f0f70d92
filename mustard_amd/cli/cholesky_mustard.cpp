// cholesky_mustard -- tiled Cholesky on one MI355X.
//
// Command line, task order and output lines of mustard/cholesky_mustard.cu
// (tiledCholesky 135-532, main 544-579); the tile operations are the
// library's gfx950 kernels instead of cuSOLVER / cuBLAS, and --subgraph runs
// the persistent device scheduler instead of device-launched CUDA graphs.
#include "cli.hpp"

int myPE = 0;

static cli::MustardConfig cfg;

static void tiledCholesky(const cli::Args &cmdl, bool subgraph)
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
        return matrix + i * B + j * B * N;  // cholesky_mustard.cu:157-160
    };

    int totalNodes = (int)T;  // cholesky_mustard.cu:200-204
    for (size_t k = 0; k < T; k++)
        for (size_t i = k + 1; i < T; i++)
            totalNodes += 2 + (int)(T - (i + 1));
    if (cfg.verbose)
        std::cout << "totalNodes=" << totalNodes << std::endl;

    mustard::TiledGraphCreator creator(s, nullptr, subgraph, totalNodes);
    const int n = (int)N, b = (int)B;
    for (size_t k = 0; k < T; k++) {
        // L[k][k] = POTRF(A[k][k])
        creator.beginCaptureOperation(std::make_pair((int)k, (int)k), {std::make_pair((int)k, (int)k)});
        checkMustard(mu_dpotrf(s, 'L', b, getMatrixBlock(d_matrix, k, k), n, d_info));
        creator.endCaptureOperation();
        for (size_t i = k + 1; i < T; i++) {
            // L[i][k] = A[i][k] L[k][k]^-T
            creator.beginCaptureOperation(std::make_pair((int)i, (int)k),
                                          {std::make_pair((int)k, (int)k), std::make_pair((int)i, (int)k)});
            checkMustard(mu_dtrsm(s, 'R', 'L', 'T', 'N', b, b, 1.0, getMatrixBlock(d_matrix, k, k),
                                  n, getMatrixBlock(d_matrix, i, k), n));
            creator.endCaptureOperation();
        }
        for (size_t i = k + 1; i < T; i++) {
            // A[i][i] -= L[i][k] L[i][k]^T (lower)
            creator.beginCaptureOperation(std::make_pair((int)i, (int)i),
                                          {std::make_pair((int)i, (int)i), std::make_pair((int)i, (int)k)});
            checkMustard(mu_dsyrk(s, 'L', 'N', b, b, -1.0, getMatrixBlock(d_matrix, i, k), n, 1.0,
                                  getMatrixBlock(d_matrix, i, i), n));
            creator.endCaptureOperation();
            for (size_t j = i + 1; j < T; j++) {
                // A[j][i] -= L[j][k] L[i][k]^T
                creator.beginCaptureOperation(
                    std::make_pair((int)j, (int)i),
                    {std::make_pair((int)j, (int)i), std::make_pair((int)j, (int)k), std::make_pair((int)i, (int)k)});
                checkMustard(mu_dgemm(s, 'N', 'T', b, b, b, -1.0, getMatrixBlock(d_matrix, j, k), n,
                                      getMatrixBlock(d_matrix, i, k), n, 1.0,
                                      getMatrixBlock(d_matrix, j, i), n));
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
    const int rc = cli::finish(cmdl, cfg, true, times, h, d_matrix, d_info, s, ex.stats().flops);
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
        cli::printSingleNodeUsage(argv[0], "Cholesky");
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
    tiledCholesky(cmdl, subgraph);
    return 0;
}
