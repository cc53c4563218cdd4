/* Prints sizeof/offsetof of the C ABI structs so tests can check the ctypes mirror. */
#include <stddef.h>
#include <stdio.h>
#include "../include/pilon_hip.h"
#include "../include/pilon_bam.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("pl_config %zu\npl_read_soa %zu\npl_call %zu\npl_region_stats %zu\npl_region_desc %zu\npl_batch_opts %zu\npl_batch_info %zu\n",
         sizeof(pl_config), sizeof(pl_read_soa), sizeof(pl_call), sizeof(pl_region_stats),
         sizeof(pl_region_desc), sizeof(pl_batch_opts), sizeof(pl_batch_info));
  P(pl_config, min_depth); P(pl_config, duplicates); P(pl_read_soa, nib); P(pl_read_soa, n_exotic);
  P(pl_region_stats, aligned_bases); P(pl_batch_info, window_ms); P(pl_call, q);
  printf("pl_serve_opts %zu\npl_serve_result %zu\n", sizeof(pl_serve_opts), sizeof(pl_serve_result));
  P(pl_serve_opts, mem_budget); P(pl_serve_opts, copy_out); P(pl_serve_result, stats);
  P(pl_serve_result, est_bytes); P(pl_serve_result, batch_device_ms);
  printf("pl_bam_stats %zu\n", sizeof(pl_bam_stats));
  P(pl_bam_stats, inflate_s); P(pl_bam_stats, threads); P(pl_bam_stats, eof_marker);
  printf("pl_asm_opts %zu\npl_asm_fix %zu\n", sizeof(pl_asm_opts), sizeof(pl_asm_fix));
  P(pl_asm_opts, min_gap); P(pl_asm_opts, fix_local); P(pl_asm_fix, ref_len); P(pl_asm_fix, patch); P(pl_asm_fix, tandem_repeat);
  return 0;
}
