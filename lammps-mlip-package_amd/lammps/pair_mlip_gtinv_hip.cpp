/* LAMMPS adapter for the HIP engine (see pair_mlip_gtinv_hip.h).
   Mirrors PairMLIPGtinv (src/USER-MLIP/pair_mlip_gtinv.cpp): flags :41-46,
   settings :545-549, coeff :555-593, init_one :599-603, compute :63-125. */
#include "pair_mlip_gtinv_hip.h"

#include <cstring>
#include <vector>

#include "atom.h"
#include "error.h"
#include "force.h"
#include "memory.h"
#include "neigh_list.h"
#include "neigh_request.h"
#include "neighbor.h"

using namespace LAMMPS_NS;

PairMLIPGtinvHIP::PairMLIPGtinvHIP(LAMMPS *lmp) : Pair(lmp) {
    single_enable = 0;
    restartinfo = 0;
    one_coeff = 1;
    manybody_flag = 1;
    // the engine returns the pair virial itself: without this flag ev_init
    // would take the F.r route (vflag_fdotr) and clear vflag_global, and
    // the virial would be lost (the Kokkos style sets it too,
    // pair_mlip_gtinv_kokkos_impl.h:278)
    no_virial_fdotr_compute = 1;
}

PairMLIPGtinvHIP::~PairMLIPGtinvHIP() {
    if (copymode) return;
    mlip_destroy(ctx);
    if (allocated) {
        memory->destroy(setflag);
        memory->destroy(cutsq);
    }
}

void PairMLIPGtinvHIP::allocate() {
    allocated = 1;
    const int n = atom->ntypes;
    memory->create(setflag, n + 1, n + 1, "pair:setflag");
    for (int i = 1; i <= n; i++)
        for (int j = i; j <= n; j++) setflag[i][j] = 0;
    memory->create(cutsq, n + 1, n + 1, "pair:cutsq");
}

void PairMLIPGtinvHIP::settings(int narg, char **) {
    force->newton_pair = 1;
    if (mlip_settings(narg) != MLIP_OK) error->all(FLERR, "Illegal pair_style command");
}

void PairMLIPGtinvHIP::coeff(int narg, char **arg) {
    if (!allocated) allocate();
    if (narg != 3 + atom->ntypes) error->all(FLERR, "Incorrect args for pair coefficients");
    if (strcmp(arg[0], "*") != 0 || strcmp(arg[1], "*") != 0)
        error->all(FLERR, "Incorrect args for pair coefficients");
    char err[1024];
    mlip_destroy(ctx);
    ctx = nullptr;
    if (mlip_create(arg[2], atom->ntypes, arg + 3, &ctx, err, sizeof err) != MLIP_OK) error->all(FLERR, err);
    std::vector<double> mass(atom->ntypes);
    mlip_masses(ctx, mass.data());
    for (int i = 1; i <= atom->ntypes; ++i) {
        if (strcmp(arg[2 + i], "NULL") != 0) atom->set_mass(FLERR, i, mass[i - 1]);
        for (int j = 1; j <= atom->ntypes; ++j) setflag[i][j] = 1;
    }
}

void PairMLIPGtinvHIP::init_style() {
    if (atom->tag_enable == 0) error->all(FLERR, "Pair style mlip_gtinv/hip requires atom IDs");
    if (force->newton_pair == 0) error->all(FLERR, "Pair style mlip_gtinv/hip requires newton pair on");
    // a full list needs no tag-based owner lookup and works with several ranks
    int irequest = neighbor->request(this, instance_me);
    neighbor->requests[irequest]->half = 0;
    neighbor->requests[irequest]->full = 1;
    full_list = 1;
}

double PairMLIPGtinvHIP::init_one(int i, int j) {
    if (setflag[i][j] == 0) error->all(FLERR, "All pair coeffs are not set");
    return mlip_cutoff(ctx);
}

void PairMLIPGtinvHIP::compute(int eflag, int vflag) {
    ev_init(eflag, vflag);
    double eng = 0.0, vir[6] = {0, 0, 0, 0, 0, 0};
    // tagint is 32-bit in a default build and 64-bit with -DLAMMPS_BIGBIG:
    // widen it to the ABI's int64 instead of reinterpreting the array
    const int nall = atom->nlocal + atom->nghost;
    tag64.resize(nall);
    for (int i = 0; i < nall; ++i) tag64[i] = static_cast<int64_t>(atom->tag[i]);
    const int rc = mlip_compute(ctx, atom->nlocal, atom->nghost, &atom->x[0][0], atom->type, tag64.data(),
                                list->inum, list->ilist, list->numneigh, list->firstneigh, full_list, eflag, 1,
                                &atom->f[0][0], &eng, vir, eflag_atom ? eatom : nullptr,
                                vflag_atom ? &vatom[0][0] : nullptr);
    if (rc != MLIP_OK) error->one(FLERR, mlip_last_error(ctx));
    if (eflag_global) eng_vdwl += eng;
    if (vflag_global)
        for (int k = 0; k < 6; ++k) virial[k] += vir[k];
}
