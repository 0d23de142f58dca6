// nbody_mtask.cpp -- T independent systems as T concurrent GPU tasks.
//
// Same flow as benchapps/nbody/utc/nbody_mtask.cc:55-275: every system has
// the same N and its own Task<BodySystemSGPU> (one worker thread and one HIP
// stream each); all run() calls are issued, then all wait() calls; the real
// wall time of the whole set is reported beside the per-task averages, and
// print_time appends {avg total, avg kernel, avg copy-in, avg copy-out, real
// total} (ms).  Options as nbody_main plus -T (number of tasks, <= 48).  The
// reference parses -O here but writes nothing; this driver writes task i's
// frames to "<-O argument>.<i>" in the frame-file format of nbody_main.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <array>
#include <memory>
#include <string>
#include <vector>

#include "Utc.h"
#include "UtcGpu.h"
#include "body_task_sgpu.h"
#include "nbody_params.h"
#include "printtime.h"
#include "task.h"

using namespace iUtc;

template <typename FTYPE>
static int run(bool printTime, int paraSelect, int iteration, int outInterval, int numBodies,
               int blocksize, MemType memtype, int ntasks, int precision, const char* outPrefix)
{
    if (numBodies % blocksize != 0)
        numBodies = (numBodies / blocksize + 1) * blocksize;
    const NBodyParams prm = select_params(paraSelect, numBodies);
    const float timestep = prm.m_timestep, softening = prm.m_softening, damping = prm.m_damping;
    const float clusterScale = prm.m_clusterScale, velocityScale = prm.m_velocityScale;
    const int nframes = iteration / outInterval + 1;
    std::vector<std::vector<FTYPE>> pos(ntasks), vel(ntasks), out(ntasks);
    Task<NbodyInit<FTYPE>> bodyInit(ProcList(0));
    for (int i = 0; i < ntasks; ++i) {
        pos[i].resize((size_t)numBodies * 4);
        vel[i].resize((size_t)numBodies * 4);
        out[i].resize((size_t)numBodies * 4 * nframes);
        bodyInit.run(NBODY_CONFIG_SHELL, pos[i].data(), vel[i].data(), (float*)nullptr, clusterScale,
                     velocityScale, numBodies, true);
        bodyInit.wait();
        std::memcpy(out[i].data(), pos[i].data(), sizeof(FTYPE) * 4 * (size_t)numBodies);
    }
    std::vector<std::unique_ptr<Task<BodySystemSGPU<FTYPE>>>> tasks;
    for (int i = 0; i < ntasks; ++i) {
        tasks.emplace_back(new Task<BodySystemSGPU<FTYPE>>(ProcList(0), TaskType::gpu_task));
        if (precision != UTCNB_PREC_REFERENCE)
            tasks.back()->exec(&BodySystemSGPU<FTYPE>::setPrecision, precision);
        tasks.back()->init((unsigned int)numBodies, (FTYPE)softening * (FTYPE)softening, (FTYPE)damping,
                           pos[i].data(), vel[i].data());
    }
    std::vector<std::array<double, 4>> rt(ntasks);
    Timer timer;
    timer.start();
    for (int i = 0; i < ntasks; ++i)
        tasks[i]->run(rt[i].data(), iteration, outInterval, blocksize, (FTYPE)timestep,
                      out[i].data() + (size_t)numBodies * 4, memtype);
    for (int i = 0; i < ntasks; ++i)
        tasks[i]->wait();
    double totaltime = timer.stop();
    if (outPrefix) {
        Task<Output<FTYPE>> outtask(ProcList(0));
        for (int i = 0; i < ntasks; ++i) {
            const std::string name = std::string(outPrefix) + "." + std::to_string(i);
            FILE* fp = std::fopen(name.c_str(), "w");
            if (!fp) {
                std::cerr << "cannot open " << name << std::endl;
                return 1;
            }
            for (int f = 0; f < nframes; ++f) {
                FTYPE timestamp = f * timestep;
                outtask.run(&fp, &out[i][(size_t)f * numBodies * 4], (float)timestamp, numBodies);
                outtask.wait();
            }
            std::fclose(fp);
        }
    }
    double runtime[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < ntasks; ++i)
        for (int j = 0; j < 4; ++j)
            runtime[j] += rt[i][j];
    std::cout << "Test complete !!!" << std::endl;
    if (printTime) {
        std::cout << "\tbody system info:" << std::endl;
        std::cout << "\t\tTotal bodies : " << numBodies << std::endl;
        std::cout << "\t\tIterations: " << iteration << std::endl;
        std::cout << "\t\tOutput frames: " << nframes << std::endl;
        std::cout << "\t\tNumber of tasks: " << ntasks << std::endl;
        std::cout << "\ttime info: " << std::endl;
        std::cout << "\t\treal total runtime: " << std::fixed << std::setprecision(4) << 1000 * totaltime << std::endl;
        std::cout << "\t\ttotal runtime: " << std::fixed << std::setprecision(4) << 1000 * runtime[0] << "(ms)" << std::endl;
        std::cout << "\t\tkernel time: " << std::fixed << std::setprecision(4) << 1000 * runtime[1] << "(ms)" << std::endl;
        std::cout << "\t\tcopyin time: " << std::fixed << std::setprecision(4) << 1000 * runtime[2] << "(ms)" << std::endl;
        std::cout << "\t\tcopyout time: " << std::fixed << std::setprecision(4) << 1000 * runtime[3] << "(ms)" << std::endl;
    }
    runtime[4] = totaltime * 1000;
    for (int j = 0; j < 4; ++j)
        runtime[j] = runtime[j] / ntasks * 1000;
    print_time(5, runtime);
    for (auto& t : tasks)
        if (t->failed())
            return 1;
    return 0;
}

int main(int argc, char** argv)
{
    bool printTime = false, useFloat = false;
    int iteration = 1, outInterval = 0, numBodies = 1024, blocksize = 256, mtype = 0, ntasks = 1;
    int paraSelect = 0;
    int precision = UTCNB_PREC_REFERENCE;
    const char* outPrefix = nullptr;
    UtcContext::getContext(argc, argv);
    int opt;
    while ((opt = getopt(argc, argv, "vt:p:m:i:l:o:O:n:b:T:P:f")) != EOF) {
        switch (opt) {
        case 'v': printTime = true; break;
        case 'm': mtype = std::atoi(optarg); break;
        case 'i': paraSelect = std::atoi(optarg); break;
        case 'l': iteration = std::atoi(optarg); break;
        case 'o': outInterval = std::atoi(optarg); break;
        case 'n': numBodies = std::atoi(optarg); break;
        case 'b': blocksize = std::atoi(optarg); break;
        case 'T': ntasks = std::atoi(optarg); break;
        case 'O': outPrefix = optarg; break;
        case 'P': precision = std::atoi(optarg); break;
        case 'f': useFloat = true; break;
        default: break;
        }
    }
    if (ntasks < 1 || ntasks > 48 || iteration < 1 || numBodies < 1 || blocksize < 64 ||
        blocksize > 1024 || blocksize % 64 != 0) {
        std::cerr << "bad arguments\n";
        return 1;
    }
    if (outInterval == 0)
        outInterval = iteration;
    MemType memtype = mtype == 1 ? MemType::pinned : (mtype == 2 ? MemType::unified : MemType::pageable);
    if (useFloat)
        return run<float>(printTime, paraSelect, iteration, outInterval, numBodies, blocksize, memtype, ntasks,
                          precision, outPrefix);
    return run<double>(printTime, paraSelect, iteration, outInterval, numBodies, blocksize, memtype, ntasks,
                       precision, outPrefix);
}
