/*
 * main.cpp - command line driver of the DAPHNE kernel interface, with the
 * reference's options and output (src/CPU/common/main.cpp:47-97):
 *   ./ndt_mapping_hip [-p N]
 * prints the elapsed kernel time and "result ok" (exit 0) or
 * "error: wrong result" (exit 1); usage errors exit 2, 3 or 4.
 */
#include <cerrno>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "benchmark.h"

static std::chrono::high_resolution_clock::time_point t_start, t_end;
static std::chrono::duration<double> elapsed(0);
static bool paused = false;
static int pipelined = 1;

extern kernel& myKernel;

static void pause_timer() {
    t_end = std::chrono::high_resolution_clock::now();
    elapsed += (t_end - t_start);
    paused = true;
}

static void unpause_timer() {
    paused = false;
    t_start = std::chrono::high_resolution_clock::now();
}

static void usage(const char* exec) {
    std::cout << "Usage: \n" << exec << " [-p N]\nOptions:\n  -p N   executes N invocations in sequence,";
    std::cout << "before taking time and check the result.\n";
    std::cout << "         Default: N=1\n";
}

int main(int argc, char** argv) {
    if ((argc != 1) && (argc != 3)) {
        usage(argv[0]);
        return 2;
    }
    if (argc == 3) {
        if (std::strcmp(argv[1], "-p") != 0) {
            usage(argv[0]);
            return 3;
        }
        errno = 0;
        pipelined = (int)std::strtol(argv[2], nullptr, 10);
        if (errno || pipelined < 1) {
            usage(argv[0]);
            return 4;
        }
        std::cout << "Invoking kernel " << pipelined << " time(s) per measure/checking step\n";
    }
    myKernel.set_timer_functions(pause_timer, unpause_timer);
    myKernel.init();
    t_start = std::chrono::high_resolution_clock::now();
    myKernel.run(pipelined);
    if (!paused) {
        t_end = std::chrono::high_resolution_clock::now();
        elapsed += t_end - t_start;
    }
    std::cout << "elapsed time: " << elapsed.count() << " seconds, average time per testcase (#"
              << myKernel.testcases << "): " << elapsed.count() / (double)myKernel.testcases << " seconds"
              << std::endl;
    if (myKernel.check_output()) {
        std::cout << "result ok\n";
        return 0;
    }
    std::cout << "error: wrong result\n";
    return 1;
}
