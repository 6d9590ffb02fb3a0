/*
 * ocm_host.h -- C entry points of the host front end (TOML readers,
 * generators, SystemGenerator), for callers that are not C++.
 *
 * ocmh_load_toml runs what the reference's read_toml_input does up to the
 * OpenMM::System (src/input/ReadTOMLInput.cpp:964-1098): parse, expand
 * includes, build the generators and the system, read the integrator and the
 * initial coordinates.  The returned handle owns the descriptors.
 *
 * ocmh_load_genesis does the same for a GENESIS `.inp` file and the `.top`
 * (with its `#include`s) and `.gro` files it names, as read_genesis_input
 * does (src/input/ReadGenesisInput.cpp:217-520).  Such inputs carry no
 * velocities: the reference starts them at zero, and ocmh_velocities returns
 * NULL.
 */
#ifndef OCM_HOST_H
#define OCM_HOST_H

#include "ocm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ocmh_input ocmh_input;

int ocmh_load_toml(const char* path, ocmh_input** out);
int ocmh_load_genesis(const char* inp_path, ocmh_input** out);
void ocmh_free(ocmh_input* in);
const ocm_system_desc* ocmh_system(ocmh_input* in);
const ocm_integrator_desc* ocmh_integrator(ocmh_input* in);
/* n*3 initial positions (nm); NULL when absent */
const double* ocmh_positions(ocmh_input* in);
/* n*3 initial velocities (nm/ps); NULL when the input gives none */
const double* ocmh_velocities(ocmh_input* in);
/* force group of the generator name, or -1 (SystemGenerator::ffname_groupid_map) */
int32_t ocmh_group_of(ocmh_input* in, const char* name);
const char* ocmh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* OCM_HOST_H */
