// nbody_mgpu.cpp -- the multi-GPU nbody driver on the gfx950 library.
//
// Same command line and flow as benchapps/nbody/utc-snode-mgpu/nbody_main.cc:63-271:
//   -v print times   -t task threads (one per GPU, at most 8)   -p processes (must be 1)
//   -m memory type (0 pageable, 1 pinned, 2 unified)   -i preset (0..6)
//   -l steps   -o output interval   -O frame file   -n bodies   -b block size
// plus -f (FP32; the reference selects FP32 or FP64 with `#define FTYPE`, the
// default here is FP64, as the reference's driver).  N is rounded up to a
// multiple of the block size, the scales follow the reference's N thresholds,
// initial conditions are the SHELL set, one Task<BodySystemMGPU> runs with
// ProcList(threads, 0), frames are written by the Output task, and print_time
// appends the thread-averaged runtime[0..3] (ms) to ./time_record.txt.  On a
// machine with fewer GPUs than threads, threads share devices.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <vector>

#include "Utc.h"
#include "UtcGpu.h"
#include "body_task_mgpu.h"
#include "nbody_params.h"
#include "printtime.h"
#include "task.h"

using namespace iUtc;

template <typename FTYPE>
static int run_nbody(bool printTime, int paraSelect, int iteration, int outInterval, int numBodies,
                     const char* outfilename, int blocksize, MemType memtype, int nthreads)
{
    if (numBodies % blocksize != 0) {
        std::cout << "number of bodies is scaled to multiple of " << blocksize << std::endl;
        numBodies = (numBodies / blocksize + 1) * blocksize;
    }
    const NBodyParams activeParams = select_params(paraSelect, numBodies);

    std::vector<FTYPE> body_pos((size_t)numBodies * 4), body_vel((size_t)numBodies * 4);
    std::vector<float> body_color((size_t)numBodies * 4);
    Task<NbodyInit<FTYPE>> bodyInit(ProcList(0));
    bodyInit.run(NBODY_CONFIG_SHELL, body_pos.data(), body_vel.data(), body_color.data(),
                 activeParams.m_clusterScale, activeParams.m_velocityScale, numBodies, true);
    bodyInit.wait();
    const int nframes = iteration / outInterval + 1;
    std::vector<FTYPE> body_outBuffer((size_t)numBodies * 4 * nframes);
    std::memcpy(body_outBuffer.data(), body_pos.data(), sizeof(FTYPE) * 4 * (size_t)numBodies);

    Task<BodySystemMGPU<FTYPE>> nbody(ProcList(nthreads, 0), TaskType::gpu_task);
    nbody.init((unsigned int)numBodies,
               (FTYPE)activeParams.m_softening * (FTYPE)activeParams.m_softening,
               (FTYPE)activeParams.m_damping, body_pos.data(), body_vel.data());
    double runtime_m[8][4] = {};
    nbody.run(runtime_m, iteration, outInterval, blocksize, (FTYPE)activeParams.m_timestep,
              body_outBuffer.data() + (size_t)numBodies * 4, memtype);
    nbody.wait();
    if (nbody.failed()) {
        std::cerr << "nbody task failed: " << nbody.error() << std::endl;
        return 1;
    }
    double runtime[4] = {0, 0, 0, 0};
    for (int t = 0; t < nthreads; ++t)
        for (int j = 0; j < 4; ++j)
            runtime[j] += runtime_m[t][j];
    for (double& r : runtime)
        r /= nthreads;

    if (outfilename) {
        FILE* fp = std::fopen(outfilename, "w");
        if (!fp) {
            std::cout << "Cann't open the output file !!!" << std::endl;
            std::exit(1);
        }
        Task<Output<FTYPE>> outtask(ProcList(0));
        for (int i = 0; i < nframes; ++i) {
            FTYPE timestamp = i * activeParams.m_timestep;
            outtask.run(&fp, &body_outBuffer[(size_t)i * numBodies * 4], (float)timestamp, numBodies);
            outtask.wait();
        }
        std::fclose(fp);
    }

    std::cout << "Test complete !!!" << std::endl;
    if (printTime) {
        std::cout << "\tbody system info:" << std::endl;
        std::cout << "\t\tTotal bodies : " << numBodies << std::endl;
        std::cout << "\t\tIterations: " << iteration << std::endl;
        std::cout << "\t\tOutput frames: " << nframes << std::endl;
        std::cout << "\t\tTask threads: " << nthreads << std::endl;
        std::cout << "\ttime info: " << std::endl;
        std::cout << "\t\ttotal runtime: " << std::fixed << std::setprecision(4) << 1000 * runtime[0] << "(ms)" << std::endl;
        std::cout << "\t\tkernel time: " << std::fixed << std::setprecision(4) << 1000 * runtime[1] << "(ms)" << std::endl;
        std::cout << "\t\tcopyin time: " << std::fixed << std::setprecision(4) << 1000 * runtime[2] << "(ms)" << std::endl;
        std::cout << "\t\tcopyout time: " << std::fixed << std::setprecision(4) << 1000 * runtime[3] << "(ms)" << std::endl;
    }
    for (double& r : runtime)
        r *= 1000;
    print_time(4, runtime);
    return 0;
}

int main(int argc, char** argv)
{
    bool printTime = false, useFloat = false;
    int paraSelect = 0, iteration = 1, outInterval = 0, numBodies = 1024, blocksize = 256;
    int nthreads = 1, nprocess = 1, mtype = 0;
    char* outfilename = nullptr;
    MemType memtype = MemType::pageable;

    UtcContext& ctx = UtcContext::getContext(argc, argv);
    int opt;
    while ((opt = getopt(argc, argv, "vt:p:m:i:l:o:O:n:b:f")) != EOF) {
        switch (opt) {
        case 'v': printTime = true; break;
        case 't': nthreads = std::atoi(optarg); break;
        case 'p': nprocess = std::atoi(optarg); break;
        case 'm': mtype = std::atoi(optarg); break;
        case 'i': paraSelect = std::atoi(optarg); break;
        case 'o': outInterval = std::atoi(optarg); break;
        case 'O': outfilename = optarg; break;
        case 'l': iteration = std::atoi(optarg); break;
        case 'n': numBodies = std::atoi(optarg); break;
        case 'b': blocksize = std::atoi(optarg); break;
        case 'f': useFloat = true; break;
        default: break;
        }
    }
    if (nprocess != ctx.numProcs()) {
        std::cerr << "process number not match with arguments '-p' !!!\n";
        return 1;
    }
    if (nthreads < 1 || nthreads > 8) {
        std::cerr << "-t must be 1..8 (the driver keeps runtime[8][4])\n";
        return 1;
    }
    if (mtype == 0)
        memtype = MemType::pageable;
    else if (mtype == 1)
        memtype = MemType::pinned;
    else if (mtype == 2)
        memtype = MemType::unified;
    else
        std::cerr << "wrong memory type for -m !!!" << std::endl;
    if (blocksize < 64 || blocksize > 1024 || blocksize % 64 != 0) {
        std::cerr << "block size must be a multiple of 64 in [64, 1024]\n";
        return 1;
    }
    if (iteration < 1 || numBodies < 1) {
        std::cerr << "need at least one step and one body\n";
        return 1;
    }
    if (outInterval == 0)
        outInterval = iteration;
    if (paraSelect > 6 || paraSelect < 0)
        paraSelect = 0;
    if (useFloat)
        return run_nbody<float>(printTime, paraSelect, iteration, outInterval, numBodies, outfilename, blocksize,
                                memtype, nthreads);
    return run_nbody<double>(printTime, paraSelect, iteration, outInterval, numBodies, outfilename, blocksize,
                             memtype, nthreads);
}
