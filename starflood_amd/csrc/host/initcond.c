/*
 * initcond.c -- initial conditions, generated on the host in FP64 and
 * written to sim.mem (and, in SF_MEM_RESIDENT mode, to the device copy).
 *
 * initcond_generate restates the reference's generator (src/initcond.c:
 * 11-263) operation for operation so that its bytes match the reference's
 * on the same libm.  The sf_ic_* generators are the seeded synthetic
 * workloads: a uniform ball and a Plummer sphere.
 */
#include <math.h>
#include <stdio.h>

#include "rng.h"
#include "sf_host.h"

#define SF_TAU 6.28318530717958647692528676655900577
#define SF_PI_4_3 4.18879020478639098461685784437267051
#define SF_PI_8_3 8.37758040957278196923371568874534102

int sf_after_host_write(const sim_t* sim_ptr, const char* func) {
	if (sf_params()->mem_mode != SF_MEM_RESIDENT) return 0;
	if (sf_sync(sim_ptr, SF_SYNC_TO_DEVICE)) {
		sf_error(func, "upload of sim.mem failed");
		return -1;
	}
	return 0;
}

static float* blk(const sim_t* sim, unsigned int e) { return (float*)sim->mem + (size_t)sim->N * e; }

/* reference src/initcond.c:11-263 */
int initcond_generate(sim_t* SF_RESTRICT sim_ptr) {
	if (!sim_ptr || !sim_ptr->mem) return STARFLOOD_FAILURE;
	const sim_t sim = *sim_ptr;
	const unsigned int N = sim.N;
	float* mas = blk(&sim, SF_E_MAS);
	float* pos = blk(&sim, SF_E_POS);
	float* vel = blk(&sim, SF_E_VEL);
	const double G = sf_params()->G;

	/* a ball of unit density: "volume" uses radius^2, harmless at R = 1 */
	const double radius = 1.000;
	const double volume = SF_PI_4_3 * (radius * radius);
	const double body_mass = volume / (double)N;
	for (unsigned int i = 0u; i < N; i++) mas[i] = (float)body_mass;
	if (sf_params()->enable_sph) {
		float* rad = blk(&sim, SF_E_RAD);
		for (unsigned int i = 0u; i < N; i++) rad[i] = (float)0.100;
	}

	for (unsigned int i = 0u; i < N; i++) {
		double p[3];
		uint32_t s[4] = {0xD88D5EFDu + (uint32_t)i, 0x203F83D3u, 0x710BE493u, 0xA294E7F1u};
		sf_pcg4d(s);
		double r[4];
		for (int j = 0; j < 4; j++) r[j] = SF_INV_U32 * (double)s[j];
		const double phi = SF_TAU * r[0], the = acos(2.0 * r[1] - 1.0);
		const double cos_phi = cos(phi), sin_phi = sin(phi);
		const double cos_the = cos(the), sin_the = sin(the);
		p[0] = cos_phi * sin_the;
		p[1] = sin_phi * sin_the;
		p[2] = cos_the;
		p[0] *= radius * cbrt(r[2]);
		p[1] *= radius * cbrt(r[2]);
		p[2] *= radius * cbrt(r[2]);
		pos[3u * i + 0u] = (float)p[0];
		pos[3u * i + 1u] = (float)p[1];
		pos[3u * i + 2u] = (float)p[2];
	}

	for (unsigned int i = 0u; i < N; i++) {
		const double p[3] = {pos[3u * i + 0u], pos[3u * i + 1u], pos[3u * i + 2u]};
		double v[3];
		uint32_t s[4] = {0x4AFDC7CFu + (uint32_t)i, 0xAE8D7CF2u, 0xE4827F00u, 0xE44CA389u};
		sf_pcg4d(s);
		double r[4];
		for (int j = 0; j < 4; j++) r[j] = SF_INV_U32 * (double)s[j];
		/* Box-Muller */
		double n[4] = {sqrt(-2.0 * log(r[0])) * cos(SF_TAU * r[1]), sqrt(-2.0 * log(r[0])) * sin(SF_TAU * r[1]),
		               sqrt(-2.0 * log(r[2])) * cos(SF_TAU * r[3]), sqrt(-2.0 * log(r[2])) * sin(SF_TAU * r[3])};
		const double sigma = 1.000e-4;
		for (int j = 0; j < 4; j++) n[j] *= sigma;
		const double velocity = sqrt((n[0] * n[0]) + (n[1] * n[1]) + (n[2] * n[2]));
		sf_pcg4d(s);
		for (int j = 0; j < 4; j++) r[j] = SF_INV_U32 * (double)s[j];
		const double phi = SF_TAU * r[0], the = acos(2.0 * r[1] - 1.0);
		const double cos_phi = cos(phi), sin_phi = sin(phi);
		const double cos_the = cos(the), sin_the = sin(the);
		v[0] = cos_phi * sin_the;
		v[1] = sin_phi * sin_the;
		v[2] = cos_the;
		for (int j = 0; j < 3; j++) v[j] *= velocity;
		/* Hubble-like expansion */
		for (int j = 0; j < 3; j++) v[j] += 4.000e-3 * sqrt(SF_PI_8_3 * G) * p[j];
		vel[3u * i + 0u] = (float)v[0];
		vel[3u * i + 1u] = (float)v[1];
		vel[3u * i + 2u] = (float)v[2];
	}
	return sf_after_host_write(sim_ptr, "initcond_generate") ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

/* Four uniforms in [0, 1) from particle i, stream k of a seed. */
static void uniforms(uint32_t seed, uint32_t i, uint32_t k, double u[4]) {
	uint32_t s[4] = {seed + i, 0x9E3779B9u * (k + 1u), 0x85EBCA6Bu ^ seed, 0xC2B2AE35u + k};
	sf_pcg4d(s);
	sf_pcg4d(s);
	for (int j = 0; j < 4; j++) u[j] = SF_INV_U32 * (double)s[j];
}

int sf_ic_uniform_ball(sim_t* sim_ptr, uint32_t seed, double radius, double total_mass, double v_hubble) {
	if (!sim_ptr || !sim_ptr->mem) return STARFLOOD_FAILURE;
	const unsigned int N = sim_ptr->N;
	float* mas = blk(sim_ptr, SF_E_MAS);
	float* pos = blk(sim_ptr, SF_E_POS);
	float* vel = blk(sim_ptr, SF_E_VEL);
	const double m = N ? total_mass / (double)N : 0.0;
	for (unsigned int i = 0; i < N; ++i) {
		double u[4];
		uniforms(seed, i, 0u, u);
		/* isotropic direction and r = R cbrt(u), as src/initcond.c:73-86 */
		const double phi = SF_TAU * u[0];
		const double ct = 2.0 * u[1] - 1.0;
		const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
		const double rr = radius * cbrt(u[2]);
		const double x[3] = {rr * cos(phi) * st, rr * sin(phi) * st, rr * ct};
		for (int k = 0; k < 3; ++k) {
			pos[3u * i + k] = (float)x[k];
			vel[3u * i + k] = (float)(v_hubble * x[k]);
		}
		mas[i] = (float)m;
	}
	return sf_after_host_write(sim_ptr, "sf_ic_uniform_ball") ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

/* Plummer sphere by the method of Aarseth, Henon and Wielen (1974, A&A 37,
   183): r from the inverted cumulative mass M(r) = M r^3 / (r^2 + a^2)^(3/2),
   speed q v_esc with q drawn by rejection from q^2 (1 - q^2)^(7/2). */
int sf_ic_plummer(sim_t* sim_ptr, uint32_t seed, unsigned int first, unsigned int count, double a, double M,
                  const double centre[3], const double bulk_v[3]) {
	if (!sim_ptr || !sim_ptr->mem || first > sim_ptr->N || count > sim_ptr->N - first || !(a > 0.0)) {
		sf_error("sf_ic_plummer", "bad arguments");
		return STARFLOOD_FAILURE;
	}
	float* mas = blk(sim_ptr, SF_E_MAS);
	float* pos = blk(sim_ptr, SF_E_POS);
	float* vel = blk(sim_ptr, SF_E_VEL);
	const double G = sf_params()->G;
	const double m = count ? M / (double)count : 0.0;
	const double c0[3] = {centre ? centre[0] : 0.0, centre ? centre[1] : 0.0, centre ? centre[2] : 0.0};
	const double b0[3] = {bulk_v ? bulk_v[0] : 0.0, bulk_v ? bulk_v[1] : 0.0, bulk_v ? bulk_v[2] : 0.0};
	for (unsigned int n = 0; n < count; ++n) {
		const unsigned int i = first + n;
		double u[4];
		uniforms(seed, i, 0u, u);
		const double X1 = fmax(u[0], 1.0e-12);
		const double r = a / sqrt(pow(X1, -2.0 / 3.0) - 1.0);
		double ct = 2.0 * u[1] - 1.0, st = sqrt(fmax(0.0, 1.0 - ct * ct)), ph = SF_TAU * u[2];
		const double x[3] = {r * st * cos(ph), r * st * sin(ph), r * ct};
		/* rejection for q; max of q^2 (1-q^2)^3.5 is 0.0920 < 0.1 */
		double q = 0.0;
		for (uint32_t k = 1u; k < 4096u; ++k) {
			double w[4];
			uniforms(seed, i, k, w);
			const double g = w[0] * w[0] * pow(1.0 - w[0] * w[0], 3.5);
			if (0.1 * w[1] < g) {
				q = w[0];
				break;
			}
			const double g2 = w[2] * w[2] * pow(1.0 - w[2] * w[2], 3.5);
			if (0.1 * w[3] < g2) {
				q = w[2];
				break;
			}
		}
		const double vesc = sqrt(2.0 * G * M / a) * pow(1.0 + r * r / (a * a), -0.25);
		const double vmag = q * vesc;
		uniforms(seed, i, 8191u, u);
		ct = 2.0 * u[0] - 1.0;
		st = sqrt(fmax(0.0, 1.0 - ct * ct));
		ph = SF_TAU * u[1];
		const double v[3] = {vmag * st * cos(ph), vmag * st * sin(ph), vmag * ct};
		for (int k = 0; k < 3; ++k) {
			pos[3u * i + k] = (float)(x[k] + c0[k]);
			vel[3u * i + k] = (float)(v[k] + b0[k]);
		}
		mas[i] = (float)m;
	}
	return sf_after_host_write(sim_ptr, "sf_ic_plummer") ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}
