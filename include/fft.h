/*
 * fft.h -- the reference's header name, forwarding to the drop-in
 * declarations in pacg_fft.h, so callers written against the reference
 * (#include "fft.h", reference fft.h:10-232) compile unchanged against
 * libpacg.so.
 */
#ifndef PACG_FFT_FORWARD_H
#define PACG_FFT_FORWARD_H
#include "pacg_fft.h"
#endif
