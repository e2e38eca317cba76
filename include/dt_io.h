/* Image reading and PNG writing over a C ABI, for bindings that do not link
 * the C++ API (no counterpart in the reference, which only has the C++
 * dt::imread / dt::imsave of dt/src/imread.cpp:9-29 and dt/src/imsave.cpp:9-30).
 * Reading takes every file the reference's imread takes (its stb_image
 * decoder with one 8-bit component: grey PNG, JPEG, PGM, TGA) with the same
 * pixels; the dt_png_ names are kept from when PNG was the only format.
 * Implemented in libdt.so.  Its error string is per thread. */
#ifndef DT_IO_H
#define DT_IO_H

#ifdef __cplusplus
extern "C" {
#endif

enum
{
  DT_IO_OK           = 0,
  DT_IO_ERR_READ     = 1, /* file missing, damaged, or of no format imread reads */
  DT_IO_ERR_FORMAT   = 2, /* readable, but not one 8-bit component; or a name without ".png" on write */
  DT_IO_ERR_WRITE    = 3,
  DT_IO_ERR_ARGUMENT = 4
};

/* Size and component count of an image file, as the reference's stb_image
 * counts components (PNG palette -> 3 or 4, tRNS adds one; JPEG 1 or 3;
 * BMP, GIF, PSD, PIC 3 or 4 and HDR 3, whose size may read as 0 since only
 * their headers are examined). */
int dt_png_info(const char* path, int* width, int* height, int* components);

/* Reads a one-component image into out (height rows of pitch bytes). */
int dt_png_read_u8(const char* path, unsigned char* out, int width, int height, long long pitch);

/* Writes 8-bit grey (channels 1) or RGB (channels 3) rows of pitch bytes. */
int dt_png_write(const char* path, const unsigned char* buf, int width, int height, long long pitch, int channels);

const char* dt_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
