/*
 * 2D subpixel smoothing helpers for TE scenes: the area fraction of a
 * cell covered by a rotated rectangle, the covered fraction of a grid
 * edge, and the permittivity mixing rules.  Same functions, arguments,
 * return codes and output conventions as the reference
 * (engine/fdtd/subpixel_te.h:1-72, subpixel_te.c:27-505).  Host code:
 * scene set-up, not the time step.
 *
 * Outputs are written only on success.  A rectangle is centred at
 * (center_x, center_y) with side size_u along the direction at angle
 * angle_rad from +x and side size_v along the perpendicular direction.
 */
#ifndef RFSYN_SUBPIXEL_TE_H
#define RFSYN_SUBPIXEL_TE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SUBPIXEL_TE_OK = 0,
    SUBPIXEL_TE_INVALID_ARGUMENT = -1,
    SUBPIXEL_TE_NUMERIC_ERROR = -2
} subpixel_te_status_t;

typedef struct {
    double center_x;
    double center_y;
    double size_u;
    double size_v;
    double angle_rad;
} subpixel_te_rectangle_t;

/* Fraction of the dx-by-dy cell centred at (sample_x, sample_y) that the
 * rectangle covers, in [0, 1] (subpixel_te.c:268-408). */
int subpixel_te_rectangle_fraction(const subpixel_te_rectangle_t *rectangle, double sample_x, double sample_y,
                                   double dx, double dy, double *fraction_out);

/* Fraction of the segment of the given length that starts at
 * (sample_x, sample_y) and runs along +x (edge_axis 1) or +y (edge_axis
 * 2) lying inside the rectangle, in the reference's parametrisation
 * (subpixel_te.c:27-86): the rectangle coordinates advance by
 * (cos a, sin a) per unit length for edge_axis 1 and by (-sin a, cos a)
 * for edge_axis 2. */
int subpixel_te_rectangle_edge_fraction(const subpixel_te_rectangle_t *rectangle, double sample_x, double sample_y,
                                        int edge_axis, double length, double *fraction_out);

/* 1 / ((1 - f) / eps_background + f / eps_inside) (subpixel_te.c:88-104). */
int subpixel_te_mix_epsilon_harmonic(double fraction, double eps_background, double eps_inside,
                                     double *epsilon_r_out);

/* Arithmetic mean (1 - f) eps_background + f eps_inside, exact at f = 0
 * and f = 1 and monotone in f (subpixel_te.c:410-455). */
int subpixel_te_mix_epsilon(double fraction, double eps_background, double eps_inside, double *epsilon_r_out);

/* subpixel_te_rectangle_fraction followed by subpixel_te_mix_epsilon
 * (subpixel_te.c:457-505); the two outputs must be distinct. */
int subpixel_te_rectangle_epsilon(const subpixel_te_rectangle_t *rectangle, double sample_x, double sample_y,
                                  double dx, double dy, double eps_background, double eps_inside,
                                  double *fraction_out, double *epsilon_r_out);

/* Smoothed 2x2 inverse permittivity for an interface with normal at
 * angle normal_rad: lam_n n n^T + lam_t (I - n n^T) with
 * lam_n = (1-f)/eps_b + f/eps_i and lam_t = 1/((1-f) eps_b + f eps_i)
 * (subpixel_te.c:106-132). */
int subpixel_te_mix_epsilon_tensor(double fraction, double eps_background, double eps_inside, double normal_rad,
                                   double *minv_xx, double *minv_xy, double *minv_yy);

#ifdef __cplusplus
}
#endif

#endif
