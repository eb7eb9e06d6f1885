/*
 * benchmark.h - the DAPHNE kernel interface (src/CPU/include/benchmark.h:11-55 of
 * the reference): a testcase-driven kernel with init / run / check_output and
 * timer callbacks that exclude file I/O from the measured time.
 */
#ifndef NDT_HARNESS_BENCHMARK_H
#define NDT_HARNESS_BENCHMARK_H

#include <cstdint>

class kernel {
public:
    /* number of testcases in the input (at least 1) */
    uint32_t testcases = 1;
    virtual ~kernel() {}
    /* reads the testcase count and prepares the device */
    virtual void init() = 0;
    /* processes all testcases, p at a time */
    virtual void run(int p = 1) = 0;
    /* compares the results with the reference data */
    virtual bool check_output() = 0;
    void set_timer_functions(void (*pause_function)(), void (*unpause_function)()) {
        unpause_func = unpause_function;
        pause_func = pause_function;
    }

protected:
    void (*unpause_func)() = nullptr;
    void (*pause_func)() = nullptr;
    /* reads up to count testcases; returns how many were read */
    virtual int read_next_testcases(int count) = 0;
};

#endif
