/*
 * Test harness over the reference's vendored PUMAS v1.2.1
 * (src/noa/3rdparty/_pumas/pumas.c), compiled in place by oracle/Makefile
 * (`make pumas-ref`) into oracle/_ref/libpumas_ref.so. It writes the golden
 * vectors of the PUMAS DCS tabulation (tests/golden/make_pumas_golden.py);
 * nothing of it is part of the library.
 *
 * The file includes pumas.c itself so that its internal routines can be
 * called: the four DCS wrappers (pumas.c:11067-11191) with the default models
 * (SSR bremsstrahlung, SSR pair production, DRSS photonuclear,
 * pumas.c:190-198) and compute_dcs_table (pumas.c:11693-11787), run on a
 * physics record that holds only what that routine reads: the kinetic
 * energy grid, the elements, the mass, the cutoff, the DCS table sizes and
 * the cross-section normalisation (set to 1, so that no row is skipped).
 */
#include "pumas.c"

static void ref_physics_init(struct pumas_physics * ph, double mass)
{
        memset(ph, 0, sizeof(*ph));
        ph->mass = mass;
        ph->dcs_bremsstrahlung = &dcs_bremsstrahlung_SSR;
        ph->dcs_pair_production = &dcs_pair_production_SSR;
        ph->dcs_photonuclear = &dcs_photonuclear_DRSS;
}

/* out[i] = dcs_get(process)(physics, element, K[i], q[i]) in m^2/kg. */
int ref_pumas_dcs(int process, double Z, double A, double I, double mass,
    const double * K, const double * q, long n, double * out)
{
        if ((process < 0) || (process > 3)) return -1;
        struct pumas_physics ph;
        ref_physics_init(&ph, mass);
        struct atomic_element el = { Z, A, I, 0, NULL };
        dcs_function_t * f = dcs_get(process);
        long i;
        for (i = 0; i < n; i++) out[i] = f(&ph, &el, K[i], q[i]);
        return 0;
}

/* The DCS table count n_table_dcs of a cutoff (pumas.c:1548-1554). */
int ref_pumas_n_table_dcs(double cutoff)
{
        int n_table_dcs = DCS_MODEL_N_REVERSE;
        if (cutoff < DCS_MODEL_X_REVERSE) {
                int n = (int)(log10(DCS_MODEL_X_REVERSE / cutoff) *
                    DCS_MODEL_PER_DECADE);
                if (n < DCS_MODEL_N_MIN) n = DCS_MODEL_N_MIN;
                n_table_dcs += n;
        }
        return n_table_dcs;
}

/*
 * compute_dcs_table for every element in turn. Outputs, in PUMAS's own
 * layouts (pumas.c:4804-4827): dcs[(p n_el + e) n_K + row][2 n_x] (log DCS
 * and PCHIP slope per node), envelope[(p n_el + e) n_K + row][2] (exponent,
 * maximum) for the processes p = 0, 1, 2, and x[n_x] (the node abscissae
 * ln(q / K), as floats).
 */
int ref_pumas_dcs_table(int n_el, const double * Z, const double * A,
    double mass, double cutoff, int n_K, const double * K, float * dcs,
    float * x, float * envelope)
{
        const int n_x = ref_pumas_n_table_dcs(cutoff);
        struct pumas_physics * ph = calloc(1, sizeof(*ph));
        struct atomic_element * els = calloc(n_el, sizeof(*els));
        struct atomic_element ** elp = calloc(n_el, sizeof(*elp));
        double * csn = malloc(sizeof(double) * N_DEL_PROCESSES * n_el * n_K);
        double * Kt = malloc(sizeof(double) * n_K);
        if (!ph || !els || !elp || !csn || !Kt) return -1;
        ref_physics_init(ph, mass);
        ph->cutoff = cutoff;
        ph->n_energies = n_K;
        ph->n_elements = n_el;
        ph->n_table_dcs = n_x;
        int i;
        for (i = 0; i < n_K; i++) Kt[i] = K[i];
        for (i = 0; i < N_DEL_PROCESSES * n_el * n_K; i++) csn[i] = 1.;
        for (i = 0; i < n_el; i++) {
                els[i].Z = Z[i];
                els[i].A = A[i];
                els[i].index = i;
                elp[i] = els + i;
        }
        ph->table_K = Kt;
        ph->table_CSn = csn;
        ph->element = elp;
        ph->table_DCS = dcs;
        ph->table_DCS_x = x;
        ph->table_DCS_envelope = envelope;
        memset(x, 0, sizeof(float) * n_x);

        struct error_context error_ = { 0 };
        int rc = 0;
        for (i = 0; i < n_el; i++) {
                if (compute_dcs_table(ph, i, &error_) != PUMAS_RETURN_SUCCESS) {
                        rc = -2;
                        break;
                }
        }
        compute_dcs_table(ph, -1, &error_);
        free(Kt);
        free(csn);
        free(elp);
        free(els);
        free(ph);
        return rc;
}
