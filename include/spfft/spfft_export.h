/* Symbol visibility for the MI355X build of the SpFFT-compatible library.
 * Replaces the CMake-generated export header of the reference build
 * (include/spfft/config.h.in:43 includes "spfft/spfft_export.h"). */
#ifndef SPFFT_EXPORT_H
#define SPFFT_EXPORT_H

#define SPFFT_EXPORT __attribute__((visibility("default")))
#define SPFFT_NO_EXPORT __attribute__((visibility("hidden")))

#endif
