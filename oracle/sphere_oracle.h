/*
 * sphere_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU (FP64, deterministic, OpenMP) restatement of the reference's DEM time
 * step and Darcy pressure solve.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the
 * timed CPU baseline; the product library never links or calls it.
 *
 * Semantics = the reference kernels in IEEE FP64 with no FMA contraction and
 * the documented deviations (contactmodel 3 Hertz-Mindlin spec, deterministic
 * bond gather, contact-slot overflow counter, r*r*r instead of pow(r,3), ...).
 */
#ifndef SPHERE_ORACLE_H_
#define SPHERE_ORACLE_H_

#include <stdint.h>
#include "../include/sphere_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

/* Copies the state (like DEM::DEM + transfer); returns NULL and fills err. */
orc *orc_create(const sph_state *s, char *err, int errlen);
void orc_destroy(orc *o);
const char *orc_error(const orc *o);
/* n loop bodies of src/device.cu:1038-2545 (output events update the same
 * host-side counters as the reference but write no files). */
int orc_step(orc *o, int n);
int orc_get_header(orc *o, sph_state *hdr);
int orc_read_state(orc *o, sph_state *out);
int orc_read_sort(orc *o, uint32_t *cellid, uint32_t *index,
                  uint32_t *cell_start, uint32_t *cell_end);
int orc_read_contacts(orc *o, uint32_t *contacts, double *distmod,
                      double *delta_t);
int orc_write_contacts(orc *o, const uint32_t *contacts, const double *distmod,
                       const double *delta_t);
int orc_darcy_stats(orc *o, sph_darcy_stats *out);
void orc_attach_ns(orc *o, void *ns);
/* (iter+1, nijac) pairs that the reference writes to <sid>-conv.log. */
int orc_conv_log(orc *o, uint32_t *pairs, int max_pairs);
/* Darcy intermediate fields with ghost layers: which = 0 p, 1 phi, 2 dphi,
 * 3 k, 4 norm, 5 dp_expl, 6 p_old; out has (nx+2)(ny+2)(nz+2) doubles. */
int orc_read_darcy_field(orc *o, int which, double *out);
void orc_set_threads(int n);
/* Contact slots per particle (32, the reference's NC, or 64) of the next
 * orc_create. */
void orc_set_contact_slots(int nc);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
