/*
 * snapshot.c -- raw snapshots of sim.mem (reference src/simulation_io.c:
 * 16-149): N * SIM_DOF reals in native byte order, no header.
 *
 * In SF_MEM_RESIDENT mode the device copy is authoritative, so sim_save_raw
 * downloads it first and sim_read_raw uploads what it read.  A short read is
 * an error here (the reference reports it and still returns success).
 */
#include <stdio.h>

#include "sf_host.h"

int sim_read_raw(sim_t* SF_RESTRICT sim_ptr, const char* SF_RESTRICT filename) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	if (!filename) {
		fprintf(stderr, "%s error: filename is NULL!\n", "sim_read_raw()");
		return STARFLOOD_FAILURE;
	}
	FILE* file = fopen(filename, "rb");
	if (!file) {
		fprintf(stderr, "%s error: fopen(\"%s\", \"%s\") ", "sim_read_raw()", filename, "rb");
		perror("failed");
		return STARFLOOD_FAILURE;
	}
	const size_t mem_length = (size_t)sim_ptr->N * (size_t)sf_sim_dof();
	int rc = 0;
	if (mem_length != fread(sim_ptr->mem, sizeof(real), mem_length, file)) {
		fprintf(stderr, "%s error: fread(file) ", "sim_read_raw()");
		perror("failed");
		rc = -1;
	}
	if (0 != fclose(file)) {
		fprintf(stderr, "%s error: fclose(file) ", "sim_read_raw()");
		perror("failed");
		rc = -1;
	}
	if (!rc && sf_after_host_write(sim_ptr, "sim_read_raw")) rc = -1;
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

int sim_save_raw(const sim_t* SF_RESTRICT sim_ptr, const char* SF_RESTRICT filename) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	if (!filename) {
		fprintf(stderr, "%s error: filename is NULL!\n", "sim_save_raw()");
		return STARFLOOD_FAILURE;
	}
	if (sf_params()->mem_mode == SF_MEM_RESIDENT && sf_sync(sim_ptr, SF_SYNC_TO_HOST)) return STARFLOOD_FAILURE;
	FILE* file = fopen(filename, "wb");
	if (!file) {
		fprintf(stderr, "%s error: fopen(\"%s\", \"%s\") ", "sim_save_raw()", filename, "wb");
		perror("failed");
		return STARFLOOD_FAILURE;
	}
	const size_t mem_length = (size_t)sim_ptr->N * (size_t)sf_sim_dof();
	int rc = 0;
	if (mem_length != fwrite(sim_ptr->mem, sizeof(real), mem_length, file)) {
		fprintf(stderr, "%s error: fwrite(file) ", "sim_save_raw()");
		perror("failed");
		rc = -1;
	}
	if (0 != fclose(file)) {
		fprintf(stderr, "%s error: fclose(file) ", "sim_save_raw()");
		perror("failed");
		rc = -1;
	}
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}
