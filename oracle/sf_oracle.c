/*
 * sf_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see sf_oracle.h).
 *
 * Built with -std=c99 -ffp-contract=off and no -ffast-math, as the
 * reference's Makefile builds it (Makefile:24,58,88), so FP32 expressions
 * round exactly as the reference's do.
 */
#define _POSIX_C_SOURCE 200112L

#include "sf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* 2^-32 rounded down to binary32 (reference src/rng.h:9) */
#define ORC_INV_U32 2.3283062977608182109179324470460414886474609375e-10
#define ORC_INV_SQRT_PI 0.56418958354775628694807945156077258584405062932899885684408572171064246844149341
#define ORC_POW_PI_3_2 5.56832799683170784528481798211883570201362439028324391075367581882974553364779570
#define ORC_PI 3.14159265358979323846264338327950288419716939937510582097494459230781640628620899

double orc_now(void) {
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return (double)t.tv_sec + 1.0e-9 * (double)t.tv_nsec;
}

int orc_num_threads(void) {
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

/* src/rng.h:13-40: LCG per lane, cross-lane multiply-adds, xorshift 16 */
void orc_pcg4d(uint32_t s[4]) {
	uint32_t v[4];
	for (int k = 0; k < 4; ++k) v[k] = s[k] * 0x0019660Du + 0x3C6EF35Fu;
	v[0] += v[1] * v[3];
	v[1] += v[2] * v[0];
	v[2] += v[0] * v[1];
	v[3] += v[1] * v[2];
	for (int k = 0; k < 4; ++k) v[k] ^= v[k] >> 16u;
	v[0] += v[1] * v[3];
	v[1] += v[2] * v[0];
	v[2] += v[0] * v[1];
	v[3] += v[1] * v[2];
	for (int k = 0; k < 4; ++k) s[k] = v[k];
}

/* Kahan step with the reference's volatile temporaries (src/gravity.c:771-784) */
static void kahan_f(float* s, float* c, float x) {
	const float y = x - *c;
	volatile float t = *s + y;
	volatile float z = t - *s;
	*c = z - y;
	*s = t;
}

/* src/gravity.c:667-819 */
void orc_pp_ref(unsigned N, const float* pos, const float* mas, double eps, double G, int kahan_acc, int kahan_pot,
                int acc_phys, float* acc, float* pot) {
	const float eps2 = (float)(eps * eps);
	const float negG = (float)(-G);
#pragma omp parallel for schedule(dynamic, 256)
	for (unsigned i = 0u; i < N; i++) {
		float as[3] = {0.0f, 0.0f, 0.0f}, ac[3] = {0.0f, 0.0f, 0.0f};
		float ps = 0.0f, pc = 0.0f;
		const float mi = mas[i];
		const float ri[3] = {pos[3u * i], pos[3u * i + 1u], pos[3u * i + 2u]};
		for (unsigned j = 0u; j < N; j++) {
			const float mj = mas[j];
			const float r[3] = {ri[0] - pos[3u * j], ri[1] - pos[3u * j + 1u], ri[2] - pos[3u * j + 2u]};
			const float r2 = (r[0] * r[0]) + (r[1] * r[1]) + (r[2] * r[2]);
			const float inv_r2 = 1.0f / (r2 + eps2);
			const float inv_r1 = 1.0f / sqrtf(r2 + eps2);
			const float pj = mj * inv_r1;
			const float F[3] = {pj * inv_r2 * r[0], pj * inv_r2 * r[1], pj * inv_r2 * r[2]};
			for (int k = 0; k < 3; k++) {
				if (kahan_acc) kahan_f(&as[k], &ac[k], F[k]);
				else as[k] += F[k];
			}
			if (kahan_pot) kahan_f(&ps, &pc, pj);
			else ps += pj;
		}
		const float f = acc_phys ? negG : negG * mi;
		for (int k = 0; k < 3; k++) acc[3u * i + k] = f * as[k];
		pot[i] = negG * ps;
	}
}

/* src/sph.c:14-283 (the density loop's unused gravity terms are omitted) */
void orc_sph_ref(unsigned N, const float* pos, const float* mas, const float* rad, double eps, int kahan,
                 float* rho, float* prs, float* acc) {
	(void)eps;
	const float isp = (float)ORC_INV_SQRT_PI;
#pragma omp parallel for schedule(dynamic, 256)
	for (unsigned i = 0u; i < N; i++) {
		float rs = 0.0f, rc = 0.0f;
		const float hi = rad[i];
		const float ri[3] = {pos[3u * i], pos[3u * i + 1u], pos[3u * i + 2u]};
		for (unsigned j = 0u; j < N; j++) {
			const float mj = mas[j];
			const float hj = rad[j];
			const float r[3] = {ri[0] - pos[3u * j], ri[1] - pos[3u * j + 1u], ri[2] - pos[3u * j + 2u]};
			const float r2 = (r[0] * r[0]) + (r[1] * r[1]) + (r[2] * r[2]);
			const float a_i = isp / hi;
			const float a_j = isp / hj;
			const float W_i = (a_i * a_i * a_i) * expf(-r2 / (hi * hi));
			const float W_j = (a_j * a_j * a_j) * expf(-r2 / (hj * hj));
			const float rho_j = 0.5f * (W_i + W_j) * mj;
			if (kahan) kahan_f(&rs, &rc, rho_j);
			else rs += rho_j;
		}
		rho[i] = rs;
	}
	/* polytrope K = 0.1, n = 1: exponent (1/n) + 1 = 2 (src/sph.c:147-159) */
	const float eos_k = 0.100f;
	/* volatile: the reference calls libm powf (whose result differs from
	   x*x in about 0.07% of cases); a constant exponent would let the
	   compiler fold the call into x*x */
	volatile float eos_exponent = (float)((1.0 / 1.0) + 1.0);
	for (unsigned i = 0u; i < N; i++) prs[i] = powf(eos_k * rho[i], eos_exponent);

	const float ipp = (float)ORC_POW_PI_3_2;
#pragma omp parallel for schedule(dynamic, 256)
	for (unsigned i = 0u; i < N; i++) {
		float as[3] = {0.0f, 0.0f, 0.0f}, ac[3] = {0.0f, 0.0f, 0.0f};
		const float mi = mas[i];
		const float hi = rad[i];
		const float ri[3] = {pos[3u * i], pos[3u * i + 1u], pos[3u * i + 2u]};
		const float rho_i = rho[i], prs_i = prs[i];
		const float rho_i2 = rho_i * rho_i;
		const float con_i = 0.0f < rho_i2 ? prs_i / rho_i2 : 0.0f;
		for (unsigned j = 0u; j < N; j++) {
			const float hj = rad[j];
			const float r[3] = {ri[0] - pos[3u * j], ri[1] - pos[3u * j + 1u], ri[2] - pos[3u * j + 2u]};
			const float rho_j = rho[j], prs_j = prs[j];
			const float rho_j2 = rho_j * rho_j;
			const float con_j = 0.0f < rho_j2 ? prs_j / rho_j2 : 0.0f;
			const float r2 = (r[0] * r[0]) + (r[1] * r[1]) + (r[2] * r[2]);
			if (r2 > (float)(1.000 * 1.000)) continue;
			const float n_i = -2.0f * expf(-r2 / (hi * hi)) / (ipp * hi * hi * hi * hi * hi);
			const float n_j = -2.0f * expf(-r2 / (hj * hj)) / (ipp * hj * hj * hj * hj * hj);
			const float g[3] = {0.5f * (n_i + n_j) * r[0], 0.5f * (n_i + n_j) * r[1], 0.5f * (n_i + n_j) * r[2]};
			const float X[3] = {(con_i + con_j) * g[0], (con_i + con_j) * g[1], (con_i + con_j) * g[2]};
			for (int k = 0; k < 3; k++) {
				if (kahan) kahan_f(&as[k], &ac[k], X[k]);
				else as[k] += X[k];
			}
		}
		for (int k = 0; k < 3; k++) acc[3u * i + k] += -mi * as[k];
	}
}

/* src/simulation.c:245-301 */
void orc_tree_bounds_ref(unsigned N, const float* pos, unsigned step_number, float bounds[6]) {
	double avg = 0.0, com = 0.0;
	for (unsigned i = 0u; i < N; i++) {
		const double p[3] = {(double)pos[3u * i], (double)pos[3u * i + 1u], (double)pos[3u * i + 2u]};
		const double r2 = (p[0] * p[0]) + (p[1] * p[1]) + (p[2] * p[2]);
		const double r1 = sqrt(r2);
		const double y = r1 - com;
		volatile double t = avg + y;
		volatile double z = t - avg;
		com = z - y;
		avg = t;
	}
	avg *= 1.0 / (double)N;
	uint32_t s[4] = {0x638FE17Bu + (uint32_t)step_number, 0xB10BDCE7u, 0x7B8ADB9Bu, 0xD8C5850Eu};
	orc_pcg4d(s);
	double j[4];
	for (int k = 0; k < 4; k++) j[k] = 2.0 * (ORC_INV_U32 * (double)s[k]) - 1.0;
	for (int i = 0; i < 3; i++) {
		bounds[i] = (float)(-2.000 * avg + 0.5 * j[i] * avg);
		bounds[3 + i] = (float)(2.000 * avg + 0.5 * j[i] * avg);
	}
}

/* src/tree.c:74-318 */
int orc_tree_build_ref(unsigned N, const float* pos, const float* mas, const float bounds[6], unsigned max_nodes,
                       int depth_max, orc_node* node, unsigned* num_nodes_out) {
	unsigned num_nodes = 1u;
	memset(&node[0], 0, sizeof(orc_node));
	for (int k = 0; k < 6; k++) node[0].bounds[k] = bounds[k];
	int rc = 0;
	for (unsigned idx = 0u; idx < N && rc == 0; idx++) {
		const float p[3] = {pos[3u * idx], pos[3u * idx + 1u], pos[3u * idx + 2u]};
		const float pm = mas[idx];
		float bmin[3] = {bounds[0], bounds[1], bounds[2]};
		float bmax[3] = {bounds[3], bounds[4], bounds[5]};
		if (bmin[0] > p[0] || p[0] > bmax[0] || bmin[1] > p[1] || p[1] > bmax[1] || bmin[2] > p[2] || p[2] > bmax[2])
			continue;
		unsigned cur = 0u;
		for (int depth = 0; depth <= depth_max; depth++) {
			const unsigned cb = node[cur].bodies;
			node[cur].bodies += 1u;
			const float mid[3] = {0.5f * (bmin[0] + bmax[0]), 0.5f * (bmin[1] + bmax[1]), 0.5f * (bmin[2] + bmax[2])};
			node[cur].params[0] += pm * p[0];
			node[cur].params[1] += pm * p[1];
			node[cur].params[2] += pm * p[2];
			node[cur].params[3] += pm;
			if (cb < 8u) {
				node[cur].branch[cb] = idx;
				break;
			}
			if (cb == 8u) {
				if (max_nodes <= num_nodes + 8u) {
					rc = -1;
					break;
				}
				if (depth_max <= depth) {
					rc = -2;
					break;
				}
				orc_node nn[8];
				memset(nn, 0, sizeof(nn));
				for (int o = 0; o < 8; o++) {
					nn[o].parent = cur;
					for (int a = 0; a < 3; a++) {
						nn[o].bounds[a] = ((o >> a) & 1) ? mid[a] : bmin[a];
						nn[o].bounds[3 + a] = ((o >> a) & 1) ? bmax[a] : mid[a];
					}
				}
				for (int k = 0; k < 8; k++) {
					const unsigned ci = node[cur].branch[k];
					const float c[3] = {pos[3u * ci], pos[3u * ci + 1u], pos[3u * ci + 2u]};
					const float cm = mas[ci];
					const int oc = (mid[0] < c[0] ? 1 : 0) + (mid[1] < c[1] ? 2 : 0) + (mid[2] < c[2] ? 4 : 0);
					const unsigned b = nn[oc].bodies;
					nn[oc].bodies += 1u;
					nn[oc].branch[b] = ci;
					nn[oc].params[0] += cm * c[0];
					nn[oc].params[1] += cm * c[1];
					nn[oc].params[2] += cm * c[2];
					nn[oc].params[3] += cm;
				}
				for (int o = 0; o < 8; o++) node[num_nodes + (unsigned)o] = nn[o];
				for (int o = 0; o < 8; o++) node[cur].branch[o] = num_nodes + (unsigned)o;
				num_nodes += 8u;
			}
			const int oct = (mid[0] <= p[0] ? 1 : 0) + (mid[1] <= p[1] ? 2 : 0) + (mid[2] <= p[2] ? 4 : 0);
			for (int a = 0; a < 3; a++) {
				bmin[a] = ((oct >> a) & 1) ? mid[a] : bmin[a];
				bmax[a] = ((oct >> a) & 1) ? bmax[a] : mid[a];
			}
			cur = node[cur].branch[oct];
		}
	}
	*num_nodes_out = num_nodes;
	if (rc) return rc;
	for (unsigned i = 0u; i < num_nodes; i++) {
		const float mass = node[i].params[3];
		const float inv = 0.0f < mass ? 1.0f / mass : 1.0f;
		node[i].params[0] *= inv;
		node[i].params[1] *= inv;
		node[i].params[2] *= inv;
	}
	return 0;
}

/* src/gravity.c:245-663 */
void orc_tree_walk_ref(unsigned N, const float* pos, const float* mas, const orc_node* node, unsigned node_cap,
                       unsigned iter_max, double eps, double G, double theta, int kahan_acc, int kahan_pot,
                       int acc_phys, float* acc, float* pot) {
	const float eps2 = (float)(eps * eps);
	const float negG = (float)(-G);
	const float th = (float)theta;
#pragma omp parallel for schedule(dynamic, 256)
	for (unsigned idx = 0u; idx < N; idx++) {
		float as[3] = {0.0f, 0.0f, 0.0f}, ac[3] = {0.0f, 0.0f, 0.0f};
		float ps = 0.0f, pc = 0.0f;
		const float ri[3] = {pos[3u * idx], pos[3u * idx + 1u], pos[3u * idx + 2u]};
		const float mi = mas[idx];
		unsigned next = 0u;
		int depth = 0;
		int path[16];
		for (int k = 0; k < 16; k++) path[k] = -1;
		for (unsigned it = 0u; it < iter_max; it++) {
			if (0 > depth) break;
			if (next >= node_cap) break;
			const orc_node nd = node[next];
			const float bmin[3] = {nd.bounds[0], nd.bounds[1], nd.bounds[2]};
			const float bmax[3] = {nd.bounds[3], nd.bounds[4], nd.bounds[5]};
			if (0u >= nd.bodies) {
				next = nd.parent;
				depth -= 1;
				continue;
			}
			if (8u < nd.bodies) {
				if (depth >= 16) break;
				const int octant = path[depth];
				if (7 <= octant) {
					next = nd.parent;
					path[depth] = -1;
					depth -= 1;
					continue;
				}
				if (0 > octant) {
					const float mid[3] = {0.5f * (bmin[0] + bmax[0]), 0.5f * (bmin[1] + bmax[1]),
					                      0.5f * (bmin[2] + bmax[2])};
					const float off[3] = {mid[0] - ri[0], mid[1] - ri[1], mid[2] - ri[2]};
					const float d2 = (off[0] * off[0]) + (off[1] * off[1]) + (off[2] * off[2]);
					const float w = (float)(1.0 / 3.0) * ((bmax[0] - bmin[0]) + (bmax[1] - bmin[1]) + (bmax[2] - bmin[2]));
					const float ratio = 0.0f < d2 ? (w * w) / d2 : 0.0f;
					if (th <= ratio) {
						next = nd.branch[0];
						path[depth] = 0;
						depth += 1;
						continue;
					}
				} else {
					next = nd.branch[octant + 1];
					path[depth] = octant + 1;
					depth += 1;
					continue;
				}
			}
			float rj[3] = {nd.params[0], nd.params[1], nd.params[2]};
			float mj = nd.params[3];
			if (bmin[0] <= ri[0] && ri[0] <= bmax[0] && bmin[1] <= ri[1] && ri[1] <= bmax[1] && bmin[2] <= ri[2] &&
			    ri[2] <= bmax[2] && 1u < nd.bodies) {
				for (int k = 0; k < 3; k++) rj[k] *= mj;
				for (int k = 0; k < 3; k++) rj[k] -= mi * ri[k];
				mj -= mi;
				for (int k = 0; k < 3; k++) rj[k] *= 0.0f < mj ? 1.0f / mj : 1.0f;
			}
			const float r[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
			const float r2 = (r[0] * r[0]) + (r[1] * r[1]) + (r[2] * r[2]);
			const float inv_r2 = 1.0f / (r2 + eps2);
			const float inv_r1 = 1.0f / sqrtf(r2 + eps2);
			const float pj = mj * inv_r1;
			const float F[3] = {pj * inv_r2 * r[0], pj * inv_r2 * r[1], pj * inv_r2 * r[2]};
			for (int k = 0; k < 3; k++) {
				if (kahan_acc) kahan_f(&as[k], &ac[k], F[k]);
				else as[k] += F[k];
			}
			if (kahan_pot) kahan_f(&ps, &pc, pj);
			else ps += pj;
			next = nd.parent;
			depth -= 1;
		}
		const float f = acc_phys ? negG : negG * mi;
		for (int k = 0; k < 3; k++) acc[3u * idx + k] = f * as[k];
		pot[idx] = negG * ps;
	}
}

static float* blk(orc_sim* s, unsigned e) { return s->mem + (size_t)s->N * e; }

/* src/simulation.c:231-345 */
int orc_sim_solv(orc_sim* s) {
	float* pos = blk(s, 0u);
	float* mas = blk(s, 9u);
	float* acc = blk(s, 6u);
	float* pot = blk(s, 10u);
	int rc = 0;
	if (s->enable_tree) {
		float b[6];
		orc_tree_bounds_ref(s->N, pos, s->step_number, b);
		s->tree_status = orc_tree_build_ref(s->N, pos, mas, b, s->tree_nodes_max, s->tree_depth_max, s->node,
		                                    &s->num_nodes);
		if (s->tree_status) rc = -1;
		orc_tree_walk_ref(s->N, pos, mas, s->node, s->tree_nodes_max, 2u * 8u * s->tree_nodes_max, s->eps, s->G,
		                  s->theta, s->kahan_solver, s->kahan_energy, s->acc_phys, acc, pot);
	} else {
		orc_pp_ref(s->N, pos, mas, s->eps, s->G, s->kahan_solver, s->kahan_energy, s->acc_phys, acc, pot);
	}
	if (s->enable_sph) orc_sph_ref(s->N, pos, mas, blk(s, 13u), s->eps, s->kahan_solver, blk(s, 14u), blk(s, 15u), acc);
	return rc;
}

/* src/simulation.c:349-519 */
int orc_sim_step(orc_sim* s) {
	const unsigned N = s->N;
	float* pos = blk(s, 0u);
	float* vel = blk(s, 3u);
	float* acc = blk(s, 6u);
	float* mas = blk(s, 9u);
	float* pot = blk(s, 10u);
	float* ken = blk(s, 11u);
	float* pen = blk(s, 12u);
	const float dt = (float)s->dt;
	for (unsigned i = 0u; i < 3u * N; i++) vel[i] += 0.5f * dt * acc[i];
	for (unsigned i = 0u; i < 3u * N; i++) pos[i] += dt * vel[i];
	const int rc = orc_sim_solv(s);
	for (unsigned i = 0u; i < 3u * N; i++) vel[i] += 0.5f * dt * acc[i];
	for (unsigned i = 0u; i < N; i++) {
		const float m = mas[i];
		const float v[3] = {vel[3u * i], vel[3u * i + 1u], vel[3u * i + 2u]};
		const float v2 = (v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]);
		ken[i] = 0.5f * m * v2;
	}
	for (unsigned i = 0u; i < N; i++) pen[i] = mas[i] * pot[i];
	double T = 0.0, Tc = 0.0, V = 0.0, Vc = 0.0;
	for (unsigned i = 0u; i < N; i++) {
		const double y = (double)ken[i] - Tc;
		volatile double t = T + y;
		volatile double z = t - T;
		Tc = z - y;
		T = t;
	}
	for (unsigned i = 0u; i < N; i++) {
		const double y = (double)pen[i] - Vc;
		volatile double t = V + y;
		volatile double z = t - V;
		Vc = z - y;
		V = t;
	}
	s->T = T;
	s->V = V;
	return rc;
}

void orc_morton63(unsigned N, const float* pos, unsigned long long* key) {
	double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
	for (unsigned i = 0u; i < N; i++)
		for (int a = 0; a < 3; a++) {
			const double v = pos[3u * i + a];
			lo[a] = v < lo[a] ? v : lo[a];
			hi[a] = v > hi[a] ? v : hi[a];
		}
	double ext = 0.0;
	for (int a = 0; a < 3; a++) ext = hi[a] - lo[a] > ext ? hi[a] - lo[a] : ext;
	if (!(ext > 0.0)) ext = 1.0;
	ext *= 1.0 + 1.0e-9;
	const double scale = 2097152.0 / ext;
	for (unsigned i = 0u; i < N; i++) {
		unsigned long long q[3], k = 0ull;
		for (int a = 0; a < 3; a++) {
			double t = ((double)pos[3u * i + a] - lo[a]) * scale;
			t = t < 0.0 ? 0.0 : (t > 2097151.0 ? 2097151.0 : t);
			q[a] = (unsigned long long)t;
		}
		for (int b = 0; b < 21; b++) {
			k |= ((q[0] >> b) & 1ull) << (3 * b + 2);
			k |= ((q[1] >> b) & 1ull) << (3 * b + 1);
			k |= ((q[2] >> b) & 1ull) << (3 * b);
		}
		key[i] = k;
	}
}

void orc_pp_f64(unsigned N, const float* pos, const float* mas, double eps, double G, int acc_phys,
                const unsigned* idx, unsigned n_idx, double* acc, double* pot) {
	const double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
	for (unsigned k = 0u; k < n_idx; k++) {
		const unsigned i = idx[k];
		const double xi = pos[3u * i], yi = pos[3u * i + 1u], zi = pos[3u * i + 2u];
		double a[3] = {0.0, 0.0, 0.0}, p = 0.0;
		for (unsigned j = 0u; j < N; j++) {
			const double rx = xi - (double)pos[3u * j], ry = yi - (double)pos[3u * j + 1u],
			             rz = zi - (double)pos[3u * j + 2u];
			const double s2 = rx * rx + ry * ry + rz * rz + eps2;
			const double inv = 1.0 / sqrt(s2);
			const double pj = (double)mas[j] * inv;
			const double w = pj * inv * inv;
			a[0] += w * rx;
			a[1] += w * ry;
			a[2] += w * rz;
			p += pj;
		}
		const double f = acc_phys ? -G : -G * (double)mas[i];
		for (int c = 0; c < 3; c++) acc[3u * k + c] = f * a[c];
		pot[k] = -G * p;
	}
}

/* M4 cubic spline with support radius h (GADGET-2 convention) */
double orc_cubic_w(double r, double h) {
	const double q = r / h, k = 8.0 / (ORC_PI * h * h * h);
	if (q <= 0.5) return k * (1.0 - 6.0 * q * q + 6.0 * q * q * q);
	if (q <= 1.0) return k * 2.0 * (1.0 - q) * (1.0 - q) * (1.0 - q);
	return 0.0;
}

double orc_cubic_dw(double r, double h) {
	const double q = r / h, k = 8.0 / (ORC_PI * h * h * h * h);
	if (q <= 0.5) return k * (-12.0 * q + 18.0 * q * q);
	if (q <= 1.0) return k * -6.0 * (1.0 - q) * (1.0 - q);
	return 0.0;
}

static void sph_cubic_density(unsigned N, const double* pos, const double* mas, const double* h, const double* u,
                              double gamma, double* rho, double* prs) {
#pragma omp parallel for schedule(dynamic, 64)
	for (unsigned i = 0u; i < N; i++) {
		double s = 0.0;
		for (unsigned j = 0u; j < N; j++) {
			const double dx = pos[3u * i] - pos[3u * j], dy = pos[3u * i + 1u] - pos[3u * j + 1u],
			             dz = pos[3u * i + 2u] - pos[3u * j + 2u];
			const double r = sqrt(dx * dx + dy * dy + dz * dz);
			s += mas[j] * 0.5 * (orc_cubic_w(r, h[i]) + orc_cubic_w(r, h[j]));
		}
		rho[i] = s;
		prs[i] = (gamma - 1.0) * s * u[i];
	}
}

/* pressure acceleration and du/dt of particle i, given every rho and P */
static void sph_cubic_force_row(unsigned N, unsigned i, const double* pos, const double* vel, const double* mas,
                                const double* h, const double* rho, const double* prs, double a[3], double* dudt) {
	double du = 0.0;
	a[0] = a[1] = a[2] = 0.0;
	const double ci = prs[i] / (rho[i] * rho[i]);
	for (unsigned j = 0u; j < N; j++) {
		if (j == i) continue;
		const double d[3] = {pos[3u * i] - pos[3u * j], pos[3u * i + 1u] - pos[3u * j + 1u],
		                     pos[3u * i + 2u] - pos[3u * j + 2u]};
		const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
		if (r <= 0.0) continue;
		const double dw = 0.5 * (orc_cubic_dw(r, h[i]) + orc_cubic_dw(r, h[j]));
		if (dw == 0.0) continue;
		const double cj = prs[j] / (rho[j] * rho[j]);
		const double f = mas[j] * (ci + cj) * dw / r;   /* grad_i Wbar = dw * d / r */
		const double vd = (vel[3u * i] - vel[3u * j]) * d[0] + (vel[3u * i + 1u] - vel[3u * j + 1u]) * d[1] +
		                  (vel[3u * i + 2u] - vel[3u * j + 2u]) * d[2];
		for (int k = 0; k < 3; k++) a[k] -= f * d[k];
		du += 0.5 * f * vd;
	}
	*dudt = du;
}

void orc_sph_cubic_f64(unsigned N, const double* pos, const double* vel, const double* mas, const double* h,
                       const double* u, double gamma, double* rho, double* prs, double* acc, double* dudt) {
	sph_cubic_density(N, pos, mas, h, u, gamma, rho, prs);
#pragma omp parallel for schedule(dynamic, 64)
	for (unsigned i = 0u; i < N; i++) sph_cubic_force_row(N, i, pos, vel, mas, h, rho, prs, acc + 3u * i, dudt + i);
}

void orc_sph_cubic_rows_f64(unsigned N, const double* pos, const double* vel, const double* mas, const double* h,
                            const double* u, double gamma, unsigned n_rows, const unsigned* rows, double* rho,
                            double* prs, double* acc_rows, double* dudt_rows) {
	sph_cubic_density(N, pos, mas, h, u, gamma, rho, prs);
#pragma omp parallel for schedule(dynamic, 16)
	for (unsigned k = 0u; k < n_rows; k++)
		sph_cubic_force_row(N, rows[k], pos, vel, mas, h, rho, prs, acc_rows + 3u * k, dudt_rows + k);
}

void orc_force_f64(unsigned N, const double* x, const double* m, double eps, double G, int acc_phys, double* acc) {
	const double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 64)
	for (unsigned i = 0u; i < N; i++) {
		double a[3] = {0.0, 0.0, 0.0};
		for (unsigned j = 0u; j < N; j++) {
			const double d[3] = {x[3u * i] - x[3u * j], x[3u * i + 1u] - x[3u * j + 1u], x[3u * i + 2u] - x[3u * j + 2u]};
			const double s2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps2;
			const double inv = 1.0 / sqrt(s2);
			const double w = m[j] * inv * inv * inv;
			for (int k = 0; k < 3; k++) a[k] += w * d[k];
		}
		const double f = acc_phys ? -G : -G * m[i];
		for (int k = 0; k < 3; k++) acc[3u * i + k] = f * a[k];
	}
}

/* acc_i = c_i sum_j m_j d_ij / s_ij^3 with c_i = -G (m_i); d acc_i / d x_k
   couples i and k through K(d) = I/s^3 - 3 d d^T / s^5:
   g_k = c_k sum_j m_j K(d_kj) mu_k - sum_i c_i m_k K(d_ik) mu_i. */
void orc_force_vjp_f64(unsigned N, const double* x, const double* m, double eps, double G, int acc_phys,
                       const double* mu, double* g) {
	const double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 64)
	for (unsigned k = 0u; k < N; k++) {
		const double ck = acc_phys ? -G : -G * m[k];
		double out[3] = {0.0, 0.0, 0.0};
		for (unsigned j = 0u; j < N; j++) {
			if (j == k) continue;
			const double d[3] = {x[3u * k] - x[3u * j], x[3u * k + 1u] - x[3u * j + 1u], x[3u * k + 2u] - x[3u * j + 2u]};
			const double s2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps2;
			const double inv = 1.0 / sqrt(s2);
			const double i3 = inv * inv * inv, i5 = i3 * inv * inv;
			const double cj = acc_phys ? -G : -G * m[j];
			/* v = ck m_j mu_k - cj m_k mu_j, then K(d) v */
			double v[3];
			for (int c = 0; c < 3; c++) v[c] = ck * m[j] * mu[3u * k + c] - cj * m[k] * mu[3u * j + c];
			const double dv = d[0] * v[0] + d[1] * v[1] + d[2] * v[2];
			for (int c = 0; c < 3; c++) out[c] += v[c] * i3 - 3.0 * d[c] * dv * i5;
		}
		for (int c = 0; c < 3; c++) g[3u * k + c] = out[c];
	}
}

void orc_kdk_f64(unsigned N, double* x, double* v, const double* m, double eps, double G, double dt, int acc_phys,
                 unsigned nsteps) {
	double* a = (double*)malloc(sizeof(double) * 3u * N);
	orc_force_f64(N, x, m, eps, G, acc_phys, a);
	for (unsigned n = 0u; n < nsteps; n++) {
		for (unsigned k = 0u; k < 3u * N; k++) v[k] += 0.5 * dt * a[k];
		for (unsigned k = 0u; k < 3u * N; k++) x[k] += dt * v[k];
		orc_force_f64(N, x, m, eps, G, acc_phys, a);
		for (unsigned k = 0u; k < 3u * N; k++) v[k] += 0.5 * dt * a[k];
	}
	free(a);
}

double orc_kdk_adjoint_f64(unsigned N, const double* x0, const double* v0, const double* m, double eps, double G,
                           double dt, int acc_phys, unsigned nsteps, const double* target, double* gx0, double* gv0) {
	const size_t n3 = 3u * (size_t)N;
	double* xs = (double*)malloc(sizeof(double) * n3 * (nsteps + 1u));
	double* v = (double*)malloc(sizeof(double) * n3);
	double* a = (double*)malloc(sizeof(double) * n3);
	double* g = (double*)malloc(sizeof(double) * n3);
	memcpy(xs, x0, sizeof(double) * n3);
	memcpy(v, v0, sizeof(double) * n3);
	orc_force_f64(N, xs, m, eps, G, acc_phys, a);
	for (unsigned n = 0u; n < nsteps; n++) {
		double* xn = xs + n3 * n;
		double* xp = xs + n3 * (n + 1u);
		for (size_t k = 0; k < n3; k++) v[k] += 0.5 * dt * a[k];
		for (size_t k = 0; k < n3; k++) xp[k] = xn[k] + dt * v[k];
		orc_force_f64(N, xp, m, eps, G, acc_phys, a);
		for (size_t k = 0; k < n3; k++) v[k] += 0.5 * dt * a[k];
	}
	const double* xN = xs + n3 * nsteps;
	double L = 0.0;
	for (size_t k = 0; k < n3; k++) L += (xN[k] - target[k]) * (xN[k] - target[k]);
	L /= (double)n3;
	/* lx = dL/dx_N, lv = dL/dv_N = 0; per step (reverse):
	   v_{n+1} = v_h + dt/2 a(x_{n+1}):  lx += dt/2 J(x_{n+1})^T lv, lvh = lv
	   x_{n+1} = x_n + dt v_h:           lvh += dt lx
	   v_h = v_n + dt/2 a(x_n):           lx += dt/2 J(x_n)^T lvh, lv = lvh */
	double* lx = gx0;
	double* lv = gv0;
	for (size_t k = 0; k < n3; k++) {
		lx[k] = 2.0 * (xN[k] - target[k]) / (double)n3;
		lv[k] = 0.0;
	}
	for (unsigned n = nsteps; n-- > 0u;) {
		double* xp = xs + n3 * (n + 1u);
		double* xn = xs + n3 * n;
		orc_force_vjp_f64(N, xp, m, eps, G, acc_phys, lv, g);
		for (size_t k = 0; k < n3; k++) lx[k] += 0.5 * dt * g[k];
		for (size_t k = 0; k < n3; k++) lv[k] += dt * lx[k];
		orc_force_vjp_f64(N, xn, m, eps, G, acc_phys, lv, g);
		for (size_t k = 0; k < n3; k++) lx[k] += 0.5 * dt * g[k];
	}
	free(xs);
	free(v);
	free(a);
	free(g);
	return L;
}
