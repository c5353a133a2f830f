// qvti_lengths.h -- the axis lengths the engine transforms, with their
// two-level split n = N1 * N2 (qvti_fft2l.h) and the build group whose
// translation unit instantiates their kernels.
//
// Every even 2/3/5-smooth length from 16 to 512 (the smooth sizes cuFFT
// plans for the reference, cuviscoelas_3D.cu:1703-1704, restricted to the
// even ones: the spectral operators' Nyquist handling needs an even n).
// N1 is the largest divisor of n not above sqrt(n) whose cofactor N2 is at
// most 32 (a pencil is N2 threads wide; the stress kernel needs at least 8
// pencils per 256-thread workgroup).  The in-register DFTs of N1 and N2 points
// are mixed-radix 4/2/3/5.
#pragma once

// X(n, N1, N2, group)
#define QVTI_LENGTHS(X)                                                                              \
    X(16, 4, 4, 0) X(18, 3, 6, 1) X(20, 4, 5, 2) X(24, 4, 6, 3) X(30, 5, 6, 4) X(32, 4, 8, 5)        \
    X(36, 6, 6, 6) X(40, 5, 8, 7) X(48, 6, 8, 0) X(50, 5, 10, 1) X(54, 6, 9, 2) X(60, 6, 10, 3)      \
    X(64, 8, 8, 4) X(72, 8, 9, 5) X(80, 8, 10, 6) X(90, 9, 10, 7) X(96, 8, 12, 0) X(100, 10, 10, 1) \
    X(108, 9, 12, 2) X(120, 10, 12, 3) X(128, 8, 16, 4) X(144, 12, 12, 5) X(150, 10, 15, 6)          \
    X(160, 10, 16, 7) X(162, 9, 18, 0) X(180, 12, 15, 1) X(192, 12, 16, 2) X(200, 10, 20, 3)        \
    X(216, 12, 18, 4) X(240, 15, 16, 5) X(250, 10, 25, 6) X(256, 16, 16, 7) X(270, 15, 18, 0)       \
    X(288, 16, 18, 1) X(300, 15, 20, 2) X(320, 16, 20, 3) X(324, 18, 18, 4) X(360, 18, 20, 5)       \
    X(384, 16, 24, 6) X(400, 20, 20, 7) X(432, 18, 24, 0) X(450, 18, 25, 1) X(480, 20, 24, 2)       \
    X(486, 18, 27, 3) X(500, 20, 25, 4) X(512, 16, 32, 5)

// Alternative splits (variant 1) of the benchmark configurations' lengths:
// the transposed order N2 x N1 where the default is not square, an unbalanced
// split where it is.  Chosen per axis by qvti_set_split or the split tuner.
#define QVTI_ALT_LENGTHS(X)                                                                           \
    X(64, 4, 16, 0) X(96, 12, 8, 1) X(100, 5, 20, 2) X(128, 16, 8, 3) X(160, 16, 10, 4)               \
    X(200, 20, 10, 5) X(256, 8, 32, 6) X(400, 16, 25, 7)

// number of build groups (translation units of qvti_kernels.hip)
#define QVTI_NGROUPS 8
