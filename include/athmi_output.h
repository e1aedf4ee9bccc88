/* athmi_output.h -- the reference's output formats written from a device-
 * resident instance (the callers after the step: data_output in main.c:226).
 *
 * The formats and file names are those of the reference's output layer:
 *   <basename>[.NNNN][.id].<ext>   (fname_construct, ath_files.c:31-63)
 *   .bin  dump_binary  (dump_binary.c:25-134): 6 ints, 4 floats, cell centres
 *         and the NVAR = 8 conserved planes as float32
 *   .ppm  output_ppm   (output_ppm.c:37-180): one expression as a P6 image
 *         through a 256-entry palette (palette.h, output.c:305-314)
 *   .tab  dump_tab     (dump_tab.c:29-233): the middle row as a text table
 * Conversions to float, the min/max scan and the palette lookup run on the
 * device; only the converted data crosses PCIe. Every file is byte-identical
 * to the reference's for the same state.
 *
 * All functions return 0 on success, an ATHMI_E* code otherwise
 * (athmi_last_error() describes it). */
#ifndef ATHMI_OUTPUT_H
#define ATHMI_OUTPUT_H

#include "athmi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* dump_binary(Grid*, Output*) (dump_binary.c:25-134) of instance inst into
 * the file path (the caller builds the name, see fname_construct). */
int athmi_dump_binary(athmi_grid *g, int inst, const char *path);

/* dump_tab(Grid*, Output*) (dump_tab.c:29-233): dat_fmt is the <outputN>
 * dat_fmt string (one floating conversion such as "%12.5e") or NULL for the
 * default " %12.8e". */
int athmi_dump_tab(athmi_grid *g, int inst, const char *path, const char *dat_fmt);

/* One <outputN> block of out_fmt = ppm (athena.h:145-183: out, palette,
 * sdmin/dmin, sdmax/dmax). expr: d, M1, M2, M3, E, B1c, B2c, B3c, ME, V1, V2,
 * V3, P, cs2, S (output.c:603-741), or the field_loop problem's user
 * expressions J3 and Bp2 (field_loop.c:160-178). palette: rainbow,
 * jh_colors, idl1, idl2, step8, step32, heat. */
typedef struct athmi_ppm_spec {
  const char *expr;
  const char *palette;
  int sdmin, sdmax; /* non-zero: scale from dmin / dmax instead of the data */
  double dmin, dmax;
} athmi_ppm_spec;

/* output_ppm(Grid*, Output*) (output_ppm.c:37-137) into the file path. num
 * is the Output's dump number: num == 0 sets *gmin, *gmax to this image's
 * data min/max, later numbers fold them in (output_ppm.c:93-100); both are
 * in/out. */
int athmi_output_ppm(athmi_grid *g, int inst, const char *path, const athmi_ppm_spec *spec,
                     int num, double *gmin, double *gmax);

#ifdef __cplusplus
}
#endif
#endif
