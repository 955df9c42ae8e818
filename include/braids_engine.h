/*
 * Reference-named operator entries of libbraids_hip (the Engine.launch
 * kernels of the river path).  Each takes DEVICE pointers, launches on the
 * stream set by braids_engine_set_stream (NULL = default stream) and returns
 * BRAIDS_OK or BRAIDS_EHIP.  Names, argument order and meaning follow the
 * reference single-source headers they replace, so the reference's
 * WaveEngine(names=...).launch(name, n, *args) binds them unchanged.
 */
#ifndef BRAIDS_ENGINE_H
#define BRAIDS_ENGINE_H
#ifdef __cplusplus
extern "C" {
#endif

int braids_engine_set_stream(void *stream);

/* src/braids_core/sediment/river_explicit.h:13-25 */
int explicit_frequency(int n, int nf, double grav, double dry, const double *q, const double *speed,
                       const double *trace, const int *mask, const int *owner, const double *normal,
                       const double *area, const double *length, const int *off, const int *faces,
                       double *out);
/* src/braids_core/sediment/river_midpoint.h:13-21 */
int midpoint_inventory(int n, double dt, const double *base, const double *area, const double *length,
                       const int *off, const int *faces, const int *signs, const double *water,
                       double *theta);
/* src/braids_core/sediment/river_midpoint.h:22-33 */
int midpoint_faces(int nf, int n, double dry, const double *q, const double *theta, const int *o,
                   const int *j, const int *bid, double *water, double *correction, int *status);
/* src/braids_core/sediment/river_midpoint.h:34-41 */
int midpoint_gather(int n, int nf, const double *area, const double *length, const int *off,
                    const int *faces, const int *signs, const double *water, const double *correction,
                    double *rate);
/* src/braids_core/sediment/river_hydraulic_drag.h:3-16 with river_chezy of
 * src/braids_core/sediment/river_roughness.py:42-47 (scalar roughness) */
int river_cell_drag(int n, double dt, double gravity, double dry, int mode, double value, int feedback,
                    const double *components, double *q, double *chezy);
/* The spatial-roughness variant generated by river_hydraulic_drag.source(True)
 * (river_hydraulic_drag.py:11-15): one roughness value per cell. */
int river_cell_drag_spatial(int n, double dt, double gravity, double dry, int mode, const double *values,
                            int feedback, const double *components, double *q, double *chezy);

#ifdef __cplusplus
}
#endif
#endif
