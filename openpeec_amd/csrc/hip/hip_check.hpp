// hip_check.hpp — HIP error handling: every runtime call is checked, and a failure ends the
// process with the call site (the library has no host fallback to hide a device error behind).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

#define HIPCHK(call)                                                                                   \
	do {                                                                                               \
		hipError_t err_ = (call);                                                                      \
		if (err_ != hipSuccess) {                                                                      \
			fprintf(stderr, "*** HIP error %s (%d) at %s:%d: %s\n", hipGetErrorString(err_), (int)err_, \
			        __FILE__, __LINE__, #call);                                                        \
			fflush(stderr);                                                                            \
			abort();                                                                                   \
		}                                                                                              \
	} while (0)
