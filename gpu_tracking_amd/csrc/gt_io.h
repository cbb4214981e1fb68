// gt_io.h -- video files read straight into caller memory in the frame's own
// sample type (u8 / u16 / f32; i8 / i16 widened to f32): multi-page TIFF
// and the Olympus ETS container (path_to_iter, src/execute_gpu.rs:1186-1270;
// decoderiter.rs:64-127 for TIFF via the tiff crate, :273-481 for ETS).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gtrack.h"

// Opens `path` by extension (tif / tiff / ets / vsi, any case); `channel`
// selects the ETS channel (default 0).  On failure *why holds the message.
gt_status gt_io_open(const char *path, int32_t has_channel, uint32_t channel, gt_file **out,
                     std::string *why);
// Frame `idx` into dst (nrows * ncols samples of the file's pipeline dtype).
// Thread-safe: concurrent reads of one file are allowed.
gt_status gt_io_read(const gt_file *f, uint64_t idx, void *dst, std::string *why);
void gt_io_close(gt_file *f);
uint64_t gt_io_nframes(const gt_file *f);
int32_t gt_io_dtype(const gt_file *f);
void gt_io_dims(const gt_file *f, uint32_t *nrows, uint32_t *ncols);
// ETS channels of the file (none for TIFF), ascending.
std::vector<uint32_t> gt_io_channels(const gt_file *f);
// Multi-page 16-bit grey TIFF, uncompressed, one strip per page (BigTIFF
// when the file would pass 4 GiB); values cast as Rust's `f32 as u16`.
gt_status gt_io_write_tiff_u16(const char *path, const float *frames, uint64_t nframes,
                               uint32_t nrows, uint32_t ncols, std::string *why);
