/*
 * test_hepers/basic.cpp -- argument parsing and result checks for the
 * self-checking skeleton programs, under the path and with the function
 * signatures of the reference (include/test_hepers/basic.cpp:41-132).
 *
 * Difference: every failed check is counted, and a program that finishes
 * with failed checks exits with status 1 (the reference prints a cross and
 * still returns success).
 */
#pragma once

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../muesli.h"

namespace msl {
namespace test_support {
inline int &failures() {
    static int f = 0;
    return f;
}
inline void exit_with_status() {
    if (failures() > 0) {
        std::fflush(nullptr);  // _Exit skips the stdio flush that exit() would do next
        std::_Exit(1);
    }
}
inline void record_failure() {
    if (failures()++ == 0) std::atexit(exit_with_status);
}
template <typename T>
inline double as_double(const T &v) {
    return static_cast<double>(v);
}
}  // namespace test_support
}  // namespace msl

inline bool to_bool(std::string str) {
    std::transform(str.begin(), str.end(), str.begin(), ::tolower);
    return str == "1" || str == "true" || str == "yes";
}

// argv: #elements, nGPUs, cpu fraction, CHECK, repetitions, threads, skeleton list
inline void arg_helper(int argc, char **argv, int &dim, int &nGPUs, int &reps, bool &CHECK, char *&skeletons) {
    msl::setNumRuns(1);
    msl::initSkeletons(argc, argv);
    dim = argc >= 2 ? std::atoi(argv[1]) : 4;
    nGPUs = argc >= 3 ? std::atoi(argv[2]) : 1;
    msl::Muesli::cpu_fraction = argc >= 4 ? std::strtod(argv[3], nullptr) : 0.0;
    CHECK = argc >= 5 ? to_bool(argv[4]) : true;
    reps = argc >= 6 ? std::atoi(argv[5]) : 1;
    msl::setNumThreads(argc >= 7 ? std::atoi(argv[6]) : 1);
    static char all[] = "all";
    skeletons = argc >= 8 ? argv[7] : all;
    if (CHECK) {
        printf("Executing with %d elements, %d repetitions, skeletons %s\n", dim, reps, skeletons);
    }
    msl::setNumGpus(nGPUs);
}

inline bool check_str_inside(const char *skeletons, const char *single_skeleton) {
    return std::strstr(skeletons, single_skeleton) != nullptr || std::strstr(skeletons, "all") != nullptr;
}

template <typename T>
void check_array_array_equal(const char *skelet, int elements, T array[], T expected_array[]) {
    for (int i = 0; i < elements; i++) {
        if (!(array[i] == expected_array[i])) {
            printf("%s \t\t\t\t \xE2\x9C\x97 At Index %d - Value %.2f != %.2f\n", skelet, i,
                   msl::test_support::as_double(array[i]), msl::test_support::as_double(expected_array[i]));
            msl::test_support::record_failure();
            return;
        }
    }
    printf("%s \t\t\t \xE2\x9C\x93\n", skelet);
}

template <typename T>
void check_array_value_equal(const char *skelet, int elements, T array[], T expected_value) {
    for (int i = 0; i < elements; i++) {
        if (!(array[i] == expected_value)) {
            printf("%s \t\t\t\t \xE2\x9C\x97 At Index %d - Value %.2f != %.2f\n", skelet, i,
                   msl::test_support::as_double(array[i]), msl::test_support::as_double(expected_value));
            msl::test_support::record_failure();
            return;
        }
    }
    printf("%s \t\t\t \xE2\x9C\x93\n", skelet);
}

template <typename T>
void check_value_value_equal(const char *skelet, T value, T expected_value) {
    if (!(value == expected_value)) {
        printf("%s \t\t\t\t \xE2\x9C\x97 Value %.2f != %.2f\n", skelet, msl::test_support::as_double(value),
               msl::test_support::as_double(expected_value));
        msl::test_support::record_failure();
    } else {
        printf("%s \t\t\t \xE2\x9C\x93\n", skelet);
    }
}

template <typename T>
void print_and_doc_runtimes(int OUTPUT, const std::string &nextfile, T runtimes[], int arraysize, int dim, int reps) {
    if (OUTPUT) {
        std::ofstream out(nextfile, std::ios_base::app);
        out << msl::Muesli::cpu_fraction << ";" << msl::Muesli::num_threads << ";" << dim << ";" << reps << ";";
        for (int i = 0; i < arraysize; i++) out << runtimes[i] << ";";
        out << "\n";
    }
    for (int i = 0; i < arraysize; i++) printf("%.4f;", msl::test_support::as_double(runtimes[i]));
    printf("\n");
}

// Four-argument form called by examples/dm_test.cpp:341 (the reference
// header only declares the six-argument form).
template <typename T>
void print_and_doc_runtimes(int OUTPUT, const std::string &nextfile, T runtimes[], int arraysize) {
    print_and_doc_runtimes(OUTPUT, nextfile, runtimes, arraysize, 0, 0);
}
