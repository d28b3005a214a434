// minifem: the host driver, mirroring the reference's main (src/main.cc:96-388) and
// FEM_loop (src/FEM.cc:139-285) at one rank, with the per-iteration work on the GPU
// through libminifem's C ABI.
//
//   minifem <mesh> <lap|ela> <iters> [options]
//     --data DIR          data root (the reference bakes DATA_PATH in at build time);
//                         default $MINIFEM_DATA or ./data
//     --schedule S        tiled (default): mf_plan_create once, then one fused
//                         mf_plan_assemble launch per iteration (assembly + prec_init +
//                         prec_inversion); colored: the drop-in calls mf_assembly_ex,
//                         mf_prec_init, mf_prec_inversion (one launch per color)
//     --precision P       fp64 | mixed; default fp64
//     --coef FILE         per-element coefficients (raw float64, file element order)
//     --store-check       write checkings/<op>_1_0 from this run instead of reading it
//     --solve RTOL        after the loop, the "solver" of the reference's main loop
//                         (src/main.cc:353, not implemented there): block-Jacobi PCG on
//                         the last iteration's operator and prec, K x = P 1 from x = 0 to
//                         RTOL (mf_pcg_ordered in the plan's node order when tiled, mf_pcg
//                         otherwise)
//   minifem gen <mesh> <lap|ela> <n> <jitter> <geomSeed> <shuffleSeed> [--data DIR]
//     writes DIR/<mesh>/inputs/<op>_1_0 for an n^3 jittered Kuhn cube
//
// Input: DIR/<mesh>/inputs/<op>_1_0 (src/IO.cc:66-68). Check:
// DIR/<mesh>/checkings/<op>_1_0 (src/IO.cc:29-31), results in numerical_results_0. The
// norms are taken on the host in the reference's serial order (compute_double_norm,
// src/FEM.cc:48-56) after the last iteration's values and prec are copied back.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "minifem.h"

namespace {

void die(const std::string &msg)
{
    std::fprintf(stderr, "Error: %s\n", msg.c_str());
    std::exit(EXIT_FAILURE);
}

void check(int st, const char *what)
{
    if (st < 0) die(std::string(what) + ": " + mf_status_string(st));
}

void hipck(hipError_t e, const char *what)
{
    if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T *dev_upload(const std::vector<T> &h)
{
    T *d = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&d), std::max<size_t>(1, h.size()) * sizeof(T)) != hipSuccess)
        die("hipMalloc failed");
    if (!h.empty() && hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        die("hipMemcpy failed");
    return d;
}

void mkdirs(const std::string &p)
{
    std::string cur;
    for (size_t i = 0; i < p.size(); i++) {
        cur += p[i];
        if (p[i] == '/' && cur.size() > 1) mkdir(cur.c_str(), 0755);
    }
    mkdir(p.c_str(), 0755);
}

std::string data_root(int argc, char **argv)
{
    for (int i = 1; i + 1 < argc; i++)
        if (!std::strcmp(argv[i], "--data")) return argv[i + 1];
    const char *env = std::getenv("MINIFEM_DATA");
    return env ? env : "data";
}

int gen_main(int argc, char **argv)
{
    if (argc < 9) die("usage: minifem gen <mesh> <lap|ela> <n> <jitter> <geomSeed> <shuffleSeed>");
    const std::string root = data_root(argc, argv), mesh = argv[2], op = argv[3];
    const int n = std::atoi(argv[4]);
    const double jitter = std::atof(argv[5]);
    const uint64_t gseed = std::strtoull(argv[6], nullptr, 0), sseed = std::strtoull(argv[7], nullptr, 0);
    int nE, nN;
    int64_t nnz;
    check(mf_cube_sizes(n, &nE, &nN, &nnz), "mf_cube_sizes");
    std::vector<double> coord((size_t)nN * 3);
    std::vector<int> e2n((size_t)nE * 4), code((size_t)nN);
    check(mf_generate_cube(n, jitter, gseed, sseed, coord.data(), e2n.data(), code.data()), "generate");
    int nb = 0;
    for (int c : code) nb += c != 0;
    const int header[6] = {nE, nN, (int)nnz, 0, 0, nb};
    const std::string dir = root + "/" + mesh + "/inputs";
    mkdirs(dir);
    const std::string path = dir + "/" + op + "_1_0";
    check(mf_store_input_data(path.c_str(), header, coord.data(), e2n.data(), nullptr, nullptr,
                              nullptr, code.data()),
          "store_input_data");
    std::printf("wrote %s: %d elements, %d nodes, %lld nnz, %d boundary nodes\n", path.c_str(),
                nE, nN, (long long)nnz, nb);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc >= 2 && !std::strcmp(argv[1], "gen")) return gen_main(argc, argv);
    if (argc < 4) {
        std::fprintf(stderr, "Please specify:\n 1. The test case.\n 2. The operator: lap or ela.\n"
                             " 3. The number of iterations.\n");
        return EXIT_FAILURE;
    }
    const std::string root = data_root(argc, argv), mesh = argv[1], opName = argv[2];
    if (opName != "lap" && opName != "ela") die("Incorrect argument \"" + opName + "\".");
    const int nbIter = (int)std::strtol(argv[3], nullptr, 0);
    if (nbIter < 1) die("Number of iterations must be at least 1.");
    std::string schedule = "tiled", precName = "fp64", coefPath;
    double solveTol = 0.0;  // > 0: solve after the loop
    bool storeCheck = false;
    for (int i = 4; i < argc; i++) {
        if (!std::strcmp(argv[i], "--schedule") && i + 1 < argc) schedule = argv[++i];
        else if (!std::strcmp(argv[i], "--precision") && i + 1 < argc) precName = argv[++i];
        else if (!std::strcmp(argv[i], "--coef") && i + 1 < argc) coefPath = argv[++i];
        else if (!std::strcmp(argv[i], "--store-check")) storeCheck = true;
        else if (!std::strcmp(argv[i], "--solve") && i + 1 < argc) {
            char *end = nullptr;
            solveTol = std::strtod(argv[++i], &end);
            if (end == argv[i] || *end || !(solveTol > 0.0))
                die(std::string("--solve needs a positive tolerance, got \"") + argv[i] + "\"");
        }
        else if (!std::strcmp(argv[i], "--data")) i++;
        else die(std::string("unknown option ") + argv[i]);
    }
    if (schedule != "tiled" && schedule != "colored")
        die("unknown schedule \"" + schedule + "\" (tiled or colored)");
    if (precName != "fp64" && precName != "mixed")
        die("unknown precision \"" + precName + "\" (fp64 or mixed)");
    const bool tiled = schedule == "tiled";
    const int opID = opName == "lap" ? MF_OP_LAP : MF_OP_ELA, dim = opID ? 9 : 1;
    const int prec = precName == "mixed" ? MF_PREC_MIXED : MF_PREC_FP64;

    std::printf("\t\t* Mini-FEM (MI355X) *\n\nTest case              : \"%s\"\n"
                "Operator               : \"%s\"\nSchedule               :  %s, %s\n"
                "Iterations             :  %d\n\n",
                mesh.c_str(), opName.c_str(), schedule.c_str(), precName.c_str(), nbIter);
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };

    // Input data (src/main.cc:140-152)
    auto t = clk::now();
    const std::string path = root + "/" + mesh + "/inputs/" + opName + "_1_0";
    int h[6];
    if (mf_read_header(path.c_str(), h) != MF_OK) die("cannot read input data: " + path);
    const int nE = h[0], nN = h[1], nbEdges = h[2], nI = h[3], nIN = h[4], nB = h[5];
    std::vector<double> coord((size_t)nN * 3);
    std::vector<int> e2n((size_t)nE * 4), nl((size_t)std::max(nI, 1) * 3), ii((size_t)nI + 1),
        in((size_t)nIN + 1), code((size_t)nN);
    check(mf_read_input_data(path.c_str(), coord.data(), e2n.data(), nl.data(), ii.data(),
                             in.data(), code.data()),
          "read_input_data");
    std::vector<double> coef;
    if (!coefPath.empty()) {
        coef.resize((size_t)nE * (opID ? 2 : 1));
        FILE *f = std::fopen(coefPath.c_str(), "rb");
        if (!f || std::fread(coef.data(), sizeof(double), coef.size(), f) != coef.size())
            die("cannot read coefficients " + coefPath);
        std::fclose(f);
    }
    std::printf("Reading input data...                done  (%.1e seconds)\n", secs(t));

    // Coloring + permutation (src/main.cc:209-235)
    t = clk::now();
    std::vector<int> perm((size_t)nE), colorToElem(MF_MAX_COLOR + 1);
    int nColors = 0;
    check(mf_coloring_creation(e2n.data(), nullptr, perm.data(), nE, nN, colorToElem.data(), &nColors),
          "coloring");
    check(mf_permute_int_2d(e2n.data(), perm.data(), nE, 4), "permute");
    if (!coef.empty()) check(mf_permute_double_2d(coef.data(), perm.data(), nE, opID ? 2 : 1), "permute coef");
    std::printf("Coloring of the mesh...              done  (%.1e seconds, %d colors)\n", secs(t), nColors);

    // CSR (src/main.cc:238-255)
    t = clk::now();
    std::vector<int> n2eI((size_t)nN + 1), n2eV((size_t)nE * 4), row((size_t)nN + 1), col((size_t)nbEdges);
    check(mf_create_nodeToElem(e2n.data(), nE, nN, n2eI.data(), n2eV.data()), "nodeToElem");
    const int nnz = mf_create_nodeToNode(row.data(), col.data(), n2eI.data(), n2eV.data(), e2n.data(), nN, nbEdges);
    check(nnz, "create_nodeToNode");
    std::printf("Creating CSR matrix...               done  (%.1e seconds)\n", secs(t));

    // Boundary conditions (src/main.cc:335-351)
    t = clk::now();
    std::vector<int> cb((size_t)nN * 3);
    check(mf_boundary_mask(code.data(), nN, nB, cb.data()), "boundary mask");
    std::printf("Computing boundary conditions...     done  (%.1e seconds)\n", secs(t));

    check(mf_device_init(0), "device init");
    double *dCoord = dev_upload(coord), *dCoef = coef.empty() ? nullptr : dev_upload(coef);
    int *dE2n = dev_upload(e2n), *dRow = dev_upload(row), *dCol = dev_upload(col), *dCb = dev_upload(cb);
    double *dValues = nullptr, *dPrec = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&dValues), (size_t)nnz * dim * sizeof(double)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&dPrec), (size_t)nN * dim * sizeof(double)) != hipSuccess)
        die("hipMalloc failed");
    // tiled: the gather schedule is built once per topology (host arrays: the permuted
    // elemToNode, the CSR pattern, the mask; coordinates group nodes into tiles)
    mf_plan *plan = nullptr;
    if (tiled) {
        t = clk::now();
        check(mf_plan_create(&plan, opID, nE, nN, nnz, e2n.data(), row.data(), col.data(),
                             cb.data(), coord.data(), nullptr),
              "mf_plan_create");
        int64_t st[8];
        check(mf_plan_stats(plan, st), "mf_plan_stats");
        std::printf("Building the tiled plan...           done  (%.1e seconds, %lld tiles)\n",
                    secs(t), (long long)st[0]);
    }
    hipStream_t stream;
    hipck(hipStreamCreate(&stream), "hipStreamCreate");
    hipEvent_t ev[5];
    for (auto &e : ev) hipck(hipEventCreate(&e), "hipEventCreate");

    std::printf("\nMain FEM loop\n");
    double phase_ms[4] = {0, 0, 0, 0};
    int timed = 0;
    for (int it = 0; it < nbIter; it++) {
        hipck(hipEventRecord(ev[0], stream), "hipEventRecord");
        if (tiled) {
            // assembly + prec_init + prec_inversion (src/FEM.cc:183-232) in one launch: the
            // whole step is reported as the assembly phase
            check(mf_plan_assemble(plan, dCoord, dCoef, dValues, dPrec, prec, nullptr, stream),
                  "mf_plan_assemble");
            for (int p = 1; p <= 4; p++) hipck(hipEventRecord(ev[p], stream), "hipEventRecord");
        } else {
            check(mf_assembly_ex(dCoord, dCoef, dValues, dRow, dCol, dE2n, nullptr, nE, nnz, dim,
                                 opID, colorToElem.data(), nColors, prec, stream),
                  "assembly");
            hipck(hipEventRecord(ev[1], stream), "hipEventRecord");
            check(mf_prec_init(dPrec, dValues, dRow, dCol, nN, dim, stream), "prec_init");
            hipck(hipEventRecord(ev[2], stream), "hipEventRecord");
            // halo exchange: nothing to exchange at one rank (src/halo.cc:44)
            hipck(hipEventRecord(ev[3], stream), "hipEventRecord");
            check(mf_prec_inversion(dPrec, dRow, dCol, dCb, nN, opID, nullptr, stream),
                  "prec_inversion");
            hipck(hipEventRecord(ev[4], stream), "hipEventRecord");
        }
        hipck(hipEventSynchronize(ev[4]), "hipEventSynchronize");
        if (nbIter == 1 || it > 0) {
            for (int p = 0; p < 4; p++) {
                float ms = 0;
                hipck(hipEventElapsedTime(&ms, ev[p], ev[p + 1]), "hipEventElapsedTime");
                phase_ms[p] += ms;
            }
            timed++;
        }
        std::printf("%d. Matrix assembly / preconditioner  done\n", it);
    }
    if (tiled) std::printf("\n(tiled: assembly and both preconditioner phases run as one fused "
                          "launch, timed as the assembly phase)\n");
    std::printf("\nAverage time (microseconds)\n----------------------------------------------\n"
                "  Matrix assembly               : %.1f\n  Preconditioner initialization : %.1f\n"
                "  Halo exchange                 : %.1f\n  Preconditioner inversion      : %.1f\n"
                "  Total                         : %.1f\n----------------------------------------------\n\n",
                1e3 * phase_ms[0] / timed, 1e3 * phase_ms[1] / timed, 1e3 * phase_ms[2] / timed,
                1e3 * phase_ms[3] / timed,
                1e3 * (phase_ms[0] + phase_ms[1] + phase_ms[2] + phase_ms[3]) / timed);

    if (solveTol > 0.0) {
        std::printf("Solving K x = P 1 (block-Jacobi PCG)...");
        std::fflush(stdout);
        const size_t n = (size_t)nN * (dim == 9 ? 3 : 1);
        double *dB = dev_upload(std::vector<double>(n, 1.0));
        double *dX = dev_upload(std::vector<double>(n, 0.0));
        int *dOrder = nullptr;
        int iters = 0;
        double rel = 0.0;
        t = clk::now();
        if (tiled) {
            hipck(hipMalloc(reinterpret_cast<void **>(&dOrder), std::max<size_t>(1, nN) * sizeof(int)),
                  "hipMalloc");
            check(mf_plan_row_order(plan, dOrder, stream), "mf_plan_row_order");
            check(mf_pcg_ordered(opID, nN, dRow, dCol, dValues, dPrec, dCb, dOrder, dB, dX, solveTol,
                                 100000, &iters, &rel, stream),
                  "mf_pcg_ordered");
        } else {
            check(mf_pcg(opID, nN, dRow, dCol, dValues, dPrec, dCb, dB, dX, solveTol, 100000, &iters,
                         &rel, stream),
                  "mf_pcg");
        }
        std::printf(" done  (%.1e seconds, %d iterations, relative residual %.3e)\n\n", secs(t),
                    iters, rel);
        for (void *p : {(void *)dB, (void *)dX, (void *)dOrder})
            if (p) hipck(hipFree(p), "hipFree");
    }

    // Numerical check (src/FEM.cc:59-98): the norms in the reference's serial order
    std::vector<double> hValues((size_t)nnz * dim), hPrec((size_t)nN * dim);
    hipck(hipMemcpy(hValues.data(), dValues, hValues.size() * sizeof(double), hipMemcpyDeviceToHost),
          "hipMemcpy");
    hipck(hipMemcpy(hPrec.data(), dPrec, hPrec.size() * sizeof(double), hipMemcpyDeviceToHost),
          "hipMemcpy");
    const double mNorm = mf_compute_double_norm(hValues.data(), (int64_t)hValues.size());
    const double pNorm = mf_compute_double_norm(hPrec.data(), (int64_t)hPrec.size());
    const std::string ckdir = root + "/" + mesh + "/checkings", ck = ckdir + "/" + opName + "_1_0";
    if (storeCheck) {
        mkdirs(ckdir);
        check(mf_store_ref_assembly(ck.c_str(), mNorm, pNorm), "store checkings");
    }
    double rM = 0, rP = 0;
    if (mf_read_ref_assembly(ck.c_str(), &rM, &rP) != MF_OK) die("cannot read reference checking: " + ck);
    char buf[1024];
    std::snprintf(buf, sizeof buf,
                  "Numerical stability of rank 0\n----------------------------------------------\n"
                  "  Matrix -> reference norm : %.6e\n              current norm : %.6e\n"
                  "                difference : %.6e\n\n    Prec -> reference norm : %.6e\n"
                  "              current norm : %.6e\n                difference : %.6e\n"
                  "----------------------------------------------\n",
                  rM, mNorm, std::fabs(rM - mNorm) / rM, rP, pNorm, std::fabs(rP - pNorm) / rP);
    std::fputs(buf, stdout);
    FILE *f = std::fopen("numerical_results_0", "w");
    if (f) { std::fputs(buf, f); std::fclose(f); }
    for (void *p : {(void *)dCoord, (void *)dCoef, (void *)dE2n, (void *)dRow, (void *)dCol,
                    (void *)dCb, (void *)dValues, (void *)dPrec})
        if (p) hipck(hipFree(p), "hipFree");
    for (auto &e : ev) hipck(hipEventDestroy(e), "hipEventDestroy");
    hipck(hipStreamDestroy(stream), "hipStreamDestroy");
    if (plan) check(mf_plan_destroy(plan), "mf_plan_destroy");
    return EXIT_SUCCESS;
}
