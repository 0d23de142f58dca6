// test_utc_shim.cpp -- CPU self-test of the UTC task shim (no GPU needed).
// Exit status 0 when every check passes; each failure prints its line.
#include <algorithm>
#include <atomic>
#include <memory>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "Utc.h"

using namespace iUtc;

static int g_fail = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);      \
            ++g_fail;                                                     \
        }                                                                 \
    } while (0)

// Records the order of jobs and the identity of the worker threads.
class Recorder : public UserTaskBase {
public:
    std::vector<std::string> log;
    std::atomic<int> runs{0}, ids_seen{0};
    int value = 0;
    std::mutex mu;
    void initImpl(int v)
    {
        if (__localThreadId == 0) {
            value = v;
            std::lock_guard<std::mutex> g(mu);
            log.push_back("init");
        }
        intra_Barrier();
        CHECK(value == v);  // every thread sees thread 0's write after the barrier
    }
    void runImpl(int* out, int k)
    {
        ids_seen.fetch_or(1 << __localThreadId);
        CHECK(__numLocalThreads > 0);
        CHECK(getCurrentTask() != nullptr);
        runs++;
        if (__localThreadId == 0) {
            std::lock_guard<std::mutex> g(mu);
            log.push_back("run" + std::to_string(k));
            out[k] = value + k;
        }
    }
    void touch(std::shared_ptr<std::vector<char>> blob)
    {
        if (__localThreadId == 0)
            value += (int)blob->size() - 1023;
    }
    void addOne()
    {
        if (__localThreadId == 0) {
            std::lock_guard<std::mutex> g(mu);
            log.push_back("exec");
            value += 1;
        }
        intra_Barrier();
    }
};

class Thrower : public UserTaskBase {
public:
    void initImpl() {}
    void runImpl(int k)
    {
        if (k == 1)
            throw std::runtime_error("boom");
    }
};

int main()
{
    UtcContext& ctx = UtcContext::getContext();
    CHECK(ctx.numProcs() == 1 && ctx.getProcRank() == 0);

    ProcList pl(4, 0);
    CHECK(pl.getNumProcs() == 4 && pl.hasProc(0) && !pl.hasProc(1));
    ProcList one(0);
    CHECK(one.getNumProcs() == 1 && one.getProc(0) == 0);

    {
        Task<Recorder> t(ProcList(4, 0));
        CHECK(t.getNumLocalThreads() == 4);
        int out[3] = {-1, -1, -1};
        t.init(10);
        t.run(out, 0);
        t.exec(&Recorder::addOne);
        t.run(out, 1);
        t.run(out, 2);
        t.wait();
        Recorder* r = t.getUserTaskObj();
        CHECK(r->runs == 12);
        CHECK(r->ids_seen == 0xF);
        CHECK(out[0] == 10 && out[1] == 12 && out[2] == 13);
        CHECK((r->log == std::vector<std::string>{"init", "run0", "exec", "run1", "run2"}));
        // wait() twice in a row and after more work
        t.wait();
        t.run(out, 0);
        t.wait();
        CHECK(r->runs == 16);
        t.finish();
        t.finish();  // idempotent
    }
    {
        // 10^5 exec jobs, each binding a 1 KiB argument: a job leaves the queue
        // (and releases its arguments) once every worker has taken it, so
        // memory stays bounded however many jobs a task runs
        Task<Recorder> t(ProcList(3, 0));
        t.init(0);
        auto blob = std::make_shared<std::vector<char>>(1024);
        size_t peak = 0;
        for (int k = 0; k < 100000; ++k) {
            t.exec(&Recorder::touch, blob);
            if (k % 1000 == 999) {
                t.wait();
                peak = std::max(peak, t.queuedJobs());
            }
        }
        t.wait();
        CHECK(t.getUserTaskObj()->value == 100000);
        CHECK(t.queuedJobs() == 0);
        CHECK(peak == 0);
        CHECK(blob.use_count() == 1);  // no job holds the bound argument any more
        t.finish();
    }
    {
        // a task with no thread on this process: every call is a no-op
        Task<Recorder> none(ProcList(2, 3));
        CHECK(none.getNumLocalThreads() == 0);
        int out[1];
        none.init(1);
        none.run(out, 0);
        none.wait();
    }
    {
        // an exception in user code is reported and never blocks wait()
        Task<Thrower> t(ProcList(2, 0));
        t.init();
        t.run(0);
        t.run(1);
        t.run(2);
        t.wait();
        CHECK(t.failed());
        CHECK(t.error() == "boom");
    }
    {
        // destructor finishes a task that was never waited for
        Task<Recorder> t(ProcList(3, 0));
        int out[3];
        t.init(1);
        t.run(out, 0);
    }
    {
        Timer tm;
        tm.start();
        std::this_thread::sleep_for(std::chrono::milliseconds(20));
        double s = tm.stop();
        CHECK(s >= 0.019 && s < 1.0);
        Timer ms(MILLISEC);
        ms.start();
        std::this_thread::sleep_for(std::chrono::milliseconds(5));
        CHECK(ms.stop() >= 4.9);
    }
    intra_Barrier();  // outside a task: no-op
    if (g_fail == 0)
        std::printf("test_utc_shim: all checks passed\n");
    return g_fail == 0 ? 0 : 1;
}
