// pl_kernels.h — launch constants shared by the kernels and the host engine.
#pragma once
#define SCAN_TILE 4096            // items per scan tile (256 threads x 16)
#define RS_ITEMS 16               // radix sort: keys per thread per block
#define RS_TILE (256 * RS_ITEMS)
#ifndef PL_WIN_BLOCK
#define PL_WIN_BLOCK 256          // threads per window workgroup
#endif
#ifndef PL_CHUNK
#define PL_CHUNK 13               // consecutive bases of one run per lane in the base loop (odd: LDS banks)
#endif
#ifndef PL_WIN_WAVES_EU
#define PL_WIN_WAVES_EU 3         // minimum waves per SIMD the window kernel is compiled for
#endif
#define PL_DEFAULT_WINDOW 512     // plain region form and autotuner fallback
#define PL_DEFAULT_BLOCK 256
#define PL_WIN_LDS_PER_POS 76     // window LDS bytes per position (counters + difference arrays)
#define PL_WIN_RECS 128           // indel evidence records a window buffers in LDS
#define PL_FIN_DELS 1024          // active deletion calls one region keeps in LDS in Pass 1
