/* oracle_capi.h -- extern "C" entry points of the CPU oracle (liboracle.so).
 * TEST INFRASTRUCTURE ONLY: for C/C++ test programs under tests/. */
#ifndef SERAC_ORACLE_CAPI_H
#define SERAC_ORACLE_CAPI_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
const char* orc_last_error(void);
void*       orc_mesh_cuboid(int nx, int ny, int nz, double lx, double ly, double lz);
int         orc_mesh_perturb(void* m, uint64_t seed, double amp);
void        orc_mesh_free(void* m);
void*       orc_space(void* m, int p, int nc, int ordering);
void        orc_space_info(void* s, int64_t* out);
void        orc_space_node_coords(void* m, void* s, double* out);
void        orc_space_free(void* s);
void*       orc_form(void* m, void* s);
void        orc_form_free(void* f);
int         orc_form_add_integral(void* f, int kind, const double* params, const int* elems, int nelem);
int         orc_form_eval(void* f, const double* U, double* R, int linearize);
int         orc_form_add_integral2(void* f, int kind, const double* params, const int* elems, int nelem,
                                   const double* state_init);
int         orc_form_eval2(void* f, const double* U, const double* U1, double* R, int linearize);
int         orc_form_jvp(void* f, const double* dU, double* dR);
int         orc_qf_eval(int kind, const double* params, int nc, const double* value, const double* grad, double* source,
                        double* flux);
int         orc_form_add_integral_state(void* f, int kind, const double* params, const int* elems, int nelem,
                                        const double* state_init);
void        orc_form_update_qdata(void* f, int flag);
int         orc_form_get_qdata(void* f, int integral, double* out);
#ifdef __cplusplus
}
#endif
#endif
