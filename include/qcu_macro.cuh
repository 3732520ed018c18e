// Constants and fatal-error helpers. The reference's convention
// (include/qcu_macro.cuh:16-40) is kept: a misuse or a runtime failure prints
// to stderr and ends the process with exit status 1.
#pragma once
#include <cstdio>
#include <cstdlib>

#include <hip/hip_runtime.h>

constexpr int Ns = 4;

#define CHECK_HIP(cmd)                                                                                  \
  do {                                                                                                  \
    hipError_t err_ = (cmd);                                                                            \
    if (err_ != hipSuccess) {                                                                           \
      fprintf(stderr, "HIP error: %s, file %s, line %d\n", hipGetErrorString(err_), __FILE__, __LINE__); \
      exit(1);                                                                                          \
    }                                                                                                   \
  } while (0)

#define errorQcu(msg)                                                            \
  do {                                                                           \
    fprintf(stderr, msg);                                                        \
    fprintf(stderr, "Error happened in file %s, line %d\n", __FILE__, __LINE__); \
    exit(1);                                                                     \
  } while (0)
