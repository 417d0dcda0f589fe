/*
 * timer.h -- wall-clock timer with the interface of the reference
 * (include/timer.h:43-107); device work is synchronised before reading it.
 */
#pragma once

#include "muesli.h"

namespace msl {

class Timer {
public:
    Timer() { start = split = end = now(); }
    double stop() {
        end = now();
        return end - start;
    }
    double splitTime() {
        double t = now();
        double r = t - split;
        split = t;
        return r;
    }
    double totalTime() { return end - start; }

private:
    static double now() {
        msl_synchronize();
        return MPI_Wtime();
    }
    double start, split, end;
};

}  // namespace msl
