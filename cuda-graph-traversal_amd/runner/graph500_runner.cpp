// graph500_runner SCALE [edgefactor] — the Graph500 driver on one MI355X.
//
// Same command line, environment flags (SKIP_BFS, SKIP_SSSP,
// SKIP_VALIDATION) and stdout block as examples/graph500_runner.cu:257-441;
// timing brackets are host wall-clock around the same calls (the reference
// uses MPI_Wtime). Additions: G500_DELTA sets the SSSP bucket width,
// G500_DUMP=1 writes the tuple dump (the reference always writes it, inside
// the generation timer, graph500_runner.cu:283), and device-time GTEPS lines
// follow the reference block.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "g500_ext.h"
#include "graph_generation.h"
#include "traversal.h"
#include "worker.h"

enum { s_minimum, s_firstquartile, s_median, s_thirdquartile, s_maximum, s_mean, s_std, s_LAST };

static constexpr int kRootCount = 64;  // graph500_runner.cu:12

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv) {
    Worker_State ws = worker_init(argc, argv, 0);
    int SCALE = 16, edgefactor = 16;
    if (argc >= 2) SCALE = atoi(argv[1]);
    if (argc >= 3) edgefactor = atoi(argv[2]);
    if (argc <= 1 || argc >= 4 || SCALE == 0 || edgefactor == 0) {
        fprintf(stderr,
                "Usage: %s SCALE edgefactor\n  SCALE = log_2(# vertices) [integer, required]\n  edgefactor = (# "
                "edges) / (# vertices) = .5 * (average vertex degree) [integer, defaults to 16]\n(Random number "
                "seed and Kronecker initiator are in graph500_runner.c)\n",
                argv[0]);
        worker_abort(&ws, 1);
    }
    const u64 seed1 = 2, seed2 = 3;
    Tuple_Graph tg = {};
    tg.nglobaledges = (u64)edgefactor << SCALE;
    u32 seed[5];
    make_mrg_seed(seed1, seed2, seed);

    const double gen_start = now();
    generate_kronecker_range(&ws, seed, (u32)SCALE, 0, (i64)tg.nglobaledges, &tg);
    if (getenv("G500_DUMP")) tuple_graph_dump(&tg, 0);
    const double gen_stop = now();
    const double gen_dev_ms = g500_last_run_ms();

    const double cons_start = now();
    Oned_Graph g = oned_graph_from_tuple_graph(&ws, &tg, (u64)1 << SCALE);
    const double cons_stop = now();
    const double cons_dev_ms = g500_last_run_ms();

    u64 roots[kRootCount];
    const int nroots = generate_bfs_roots(&ws, &g, seed1, seed2, roots, kRootCount);
    std::vector<double> stats(s_LAST);

    if (!getenv("SKIP_BFS") && nroots > 0) {
        Bfs_State bfs = oned_graph_bfs_create(&ws, &g);
        std::vector<double> times(nroots), vtimes(nroots), nedge(nroots), dev_ms(nroots);
        int completed = 0;
        bool all_valid = true;
        oned_graph_bfs_clear(&ws, &bfs);
        oned_graph_bfs_run(&ws, &g, &bfs, roots[0]);  // warm-up, like the reference
        for (int i = 0; i < nroots; ++i) {
            oned_graph_bfs_clear(&ws, &bfs);
            const double t0 = now();
            oned_graph_bfs_run(&ws, &g, &bfs, roots[i]);
            const double t1 = now();
            times[i] = t1 - t0;
            dev_ms[i] = g500_last_run_ms();
            nedge[i] = (double)oned_graph_edge_count(&ws, &g, bfs.pred);
            if (!getenv("SKIP_VALIDATION")) {
                const double v0 = now();
                const bool ok = oned_graph_bfs_validate(&ws, &g, &bfs, roots[i], nullptr);
                vtimes[i] = now() - v0;
                if (!ok) {
                    fprintf(stderr, "Validation failed for this BFS root; skipping rest.\n");
                    all_valid = false;
                    break;
                }
            } else {
                vtimes[i] = -1.0;
            }
            completed = i + 1;
        }
        if (all_valid && completed > 0) {
            double *r = stats.data();
            fprintf(stdout, "SCALE:                          %d\n", SCALE);
            fprintf(stdout, "edgefactor:                     %d\n", edgefactor);
            fprintf(stdout, "NBFS:                           %d\n", completed);
            fprintf(stdout, "graph_generation:               %g\n", gen_stop - gen_start);
            fprintf(stdout, "num_mpi_processes:              %d\n", ws.rank_size);
            fprintf(stdout, "construction_time:              %g\n", cons_stop - cons_start);
            get_statistics(times.data(), completed, r);
            fprintf(stdout, "bfs  min_time:                  %g\n", r[s_minimum]);
            fprintf(stdout, "bfs  firstquartile_time:        %g\n", r[s_firstquartile]);
            fprintf(stdout, "bfs  median_time:               %g\n", r[s_median]);
            fprintf(stdout, "bfs  thirdquartile_time:        %g\n", r[s_thirdquartile]);
            fprintf(stdout, "bfs  max_time:                  %g\n", r[s_maximum]);
            fprintf(stdout, "bfs  mean_time:                 %g\n", r[s_mean]);
            fprintf(stdout, "bfs  stddev_time:               %g\n", r[s_std]);
            get_statistics(nedge.data(), completed, r);
            fprintf(stdout, "min_nedge:                      %.11g\n", r[s_minimum]);
            fprintf(stdout, "firstquartile_nedge:            %.11g\n", r[s_firstquartile]);
            fprintf(stdout, "median_nedge:                   %.11g\n", r[s_median]);
            fprintf(stdout, "thirdquartile_nedge:            %.11g\n", r[s_thirdquartile]);
            fprintf(stdout, "max_nedge:                      %.11g\n", r[s_maximum]);
            fprintf(stdout, "mean_nedge:                     %.11g\n", r[s_mean]);
            fprintf(stdout, "stddev_nedge:                   %.11g\n", r[s_std]);
            std::vector<double> spe(completed);
            for (int i = 0; i < completed; ++i) spe[i] = times[i] / nedge[i];
            get_statistics(spe.data(), completed, r);
            fprintf(stdout, "bfs  min_TEPS:                  %g\n", 1.0 / r[s_maximum]);
            fprintf(stdout, "bfs  firstquartile_TEPS:        %g\n", 1.0 / r[s_thirdquartile]);
            fprintf(stdout, "bfs  median_TEPS:               %g\n", 1.0 / r[s_median]);
            fprintf(stdout, "bfs  thirdquartile_TEPS:        %g\n", 1.0 / r[s_firstquartile]);
            fprintf(stdout, "bfs  max_TEPS:                  %g\n", 1.0 / r[s_minimum]);
            fprintf(stdout, "bfs  harmonic_mean_TEPS:     !  %g\n", 1.0 / r[s_mean]);
            fprintf(stdout, "bfs  harmonic_stddev_TEPS:      %g\n",
                    r[s_std] / (r[s_mean] * r[s_mean] * std::sqrt((double)(completed - 1))));
            if (!getenv("SKIP_VALIDATION")) {
                get_statistics(vtimes.data(), completed, r);
                fprintf(stdout, "bfs  min_validate:              %g\n", r[s_minimum]);
                fprintf(stdout, "bfs  firstquartile_validate:    %g\n", r[s_firstquartile]);
                fprintf(stdout, "bfs  median_validate:           %g\n", r[s_median]);
                fprintf(stdout, "bfs  thirdquartile_validate:    %g\n", r[s_thirdquartile]);
                fprintf(stdout, "bfs  max_validate:              %g\n", r[s_maximum]);
                fprintf(stdout, "bfs  mean_validate:             %g\n", r[s_mean]);
                fprintf(stdout, "bfs  stddev_validate:           %g\n", r[s_std]);
            }
            double hsum = 0;
            for (int i = 0; i < completed; ++i) hsum += dev_ms[i] * 1e-3 / nedge[i];
            fprintf(stdout, "bfs  harmonic_mean_GTEPS_device: %g\n", completed / hsum * 1e-9);
            fprintf(stdout, "graph_generation_device_ms:     %g\n", gen_dev_ms);
            fprintf(stdout, "construction_device_ms:         %g\n", cons_dev_ms);
        } else if (!all_valid) {
            fprintf(stdout, "No results printed for invalid run.\n");
        }
        oned_graph_bfs_destroy(&bfs);
    }

    if (!getenv("SKIP_SSSP") && nroots > 0) {
        Sssp_State sssp = oned_graph_sssp_create(&ws, &g);
        if (const char *d = getenv("G500_DELTA")) sssp.delta = (f32)atof(d);
        std::vector<double> times(nroots), nedge(nroots), dev_ms(nroots);
        oned_graph_sssp_clear(&ws, &sssp);
        oned_graph_sssp_run(&ws, &g, &sssp, roots[0]);
        for (int i = 0; i < nroots; ++i) {
            oned_graph_sssp_clear(&ws, &sssp);
            const double t0 = now();
            oned_graph_sssp_run(&ws, &g, &sssp, roots[i]);
            times[i] = now() - t0;
            dev_ms[i] = g500_last_run_ms();
            nedge[i] = (double)oned_graph_edge_count(&ws, &g, sssp.pred);
        }
        double *r = stats.data();
        get_statistics(times.data(), nroots, r);
        fprintf(stdout, "sssp min_time:                  %g\n", r[s_minimum]);
        fprintf(stdout, "sssp firstquartile_time:        %g\n", r[s_firstquartile]);
        fprintf(stdout, "sssp median_time:               %g\n", r[s_median]);
        fprintf(stdout, "sssp thirdquartile_time:        %g\n", r[s_thirdquartile]);
        fprintf(stdout, "sssp max_time:                  %g\n", r[s_maximum]);
        fprintf(stdout, "sssp mean_time:                 %g\n", r[s_mean]);
        fprintf(stdout, "sssp stddev_time:               %g\n", r[s_std]);
        std::vector<double> spe(nroots);
        for (int i = 0; i < nroots; ++i) spe[i] = times[i] / nedge[i];
        get_statistics(spe.data(), nroots, r);
        fprintf(stdout, "sssp min_TEPS:                  %g\n", 1.0 / r[s_maximum]);
        fprintf(stdout, "sssp median_TEPS:               %g\n", 1.0 / r[s_median]);
        fprintf(stdout, "sssp max_TEPS:                  %g\n", 1.0 / r[s_minimum]);
        double hsum = 0;
        for (int i = 0; i < nroots; ++i) hsum += dev_ms[i] * 1e-3 / nedge[i];
        fprintf(stdout, "sssp harmonic_mean_GTEPS_device: %g\n", nroots / hsum * 1e-9);
        oned_graph_sssp_destroy(&sssp);
    }
    oned_graph_free(&g);
    worker_finalize(&ws);
    return 0;
}
