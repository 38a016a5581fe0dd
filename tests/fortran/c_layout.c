/* Layout of the C structs of coral_amd.h, in the format of binding_layout.f90,
 * so that the tests compare the Fortran bind(C) types with the C compiler's
 * own layout. */
#include <stddef.h>
#include <stdio.h>

#include "coral_amd.h"

#define SZ(T) printf(#T " %zu\n", sizeof(T))
#define OFF(T, F) printf(#T "." #F " %zu\n", offsetof(T, F))

int main(void) {
    SZ(coral_amd_piece);
    SZ(coral_amd_nlterm);
    SZ(coral_amd_system);
    SZ(coral_amd_fullterm);
    SZ(coral_amd_fullvar);
    SZ(coral_amd_quadvar);
    SZ(coral_amd_problem);
    OFF(coral_amd_piece, eqn);
    OFF(coral_amd_nlterm, qvar);
    OFF(coral_amd_system, bc_codes);
    OFF(coral_amd_system, mass);
    OFF(coral_amd_system, stif);
    OFF(coral_amd_system, nl);
    OFF(coral_amd_fullterm, position);
    OFF(coral_amd_fullvar, terms);
    OFF(coral_amd_problem, Lx);
    OFF(coral_amd_problem, nstages);
    OFF(coral_amd_problem, a_arr);
    OFF(coral_amd_problem, kxky);
    OFF(coral_amd_problem, zero);
    OFF(coral_amd_problem, full);
    OFF(coral_amd_problem, quad);
    return 0;
}
