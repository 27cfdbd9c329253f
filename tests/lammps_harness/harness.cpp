/* Drives a LAMMPS pair style through the calls a LAMMPS run makes, on the
   stand-in LAMMPS classes of this directory (test harness only):

     pair_style / pair_coeff  ->  settings(0), coeff(* * potfile E1 E2 ...)
     init                     ->  init_style() (the list request), init_one(i, j)
     each step                ->  force_clear (f of owned atoms and ghosts),
                                  compute(eflag, vflag)

   usage: lammps_harness <style> <potfile> <in.bin> <out.bin> <E1> [E2 ...]
     style: mlip_gtinv/hip (the HIP adapter) or fdotr_probe (a style
     without no_virial_fdotr_compute, which shows what LAMMPS's virial
     flags do to a style that leaves virial_fdotr_compute uncalled)
   in.bin (native endianness): int32 nlocal, nghost, ntypes, inum, nnbr,
     eflag, vflag, nsteps; float64 x[nall][3]; int32 type[nall];
     int32 tag[nall]; int32 ilist[inum]; int32 numneigh[nlocal];
     int32 neighbors[nnbr] (the lists of ilist[0], ilist[1], ... in order)
   out.bin: int32 request_full, request_half, newton_pair, vflag_global,
     vflag_fdotr, eflag_atom, vflag_atom, pad; float64 cutoff,
     mass[ntypes], then per step: eng_vdwl, virial[6], f[nall][3],
     eatom[nall], vatom[nall][6] (zeros where the flag is off). */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "atom.h"
#include "error.h"
#include "force.h"
#include "memory.h"
#include "neigh_list.h"
#include "neigh_request.h"
#include "neighbor.h"
#include "pair.h"
#include "pair_mlip_gtinv_hip.h"

using namespace LAMMPS_NS;

// ------------------------------------------------------------ stand-in classes
LAMMPS::LAMMPS() {
    memory = new Memory(this);
    error = new Error(this);
    atom = new Atom(this);
    force = new Force(this);
    neighbor = new Neighbor(this);
}

LAMMPS::~LAMMPS() {
    delete neighbor;
    delete force;
    delete atom;
    delete error;
    delete memory;
}

Atom::~Atom() {
    memory->destroy(x);
    memory->destroy(f);
    memory->destroy(type);
    memory->destroy(tag);
    memory->destroy(mass);
    memory->destroy(mass_setflag);
}

void Atom::grow(int n, int ntype) {
    nmax = n;
    ntypes = ntype;
    memory->create(x, n, 3, "atom:x");
    memory->create(f, n, 3, "atom:f");
    memory->create(type, n, "atom:type");
    memory->create(tag, n, "atom:tag");
    memory->create(mass, ntype + 1, "atom:mass");
    memory->create(mass_setflag, ntype + 1, "atom:mass_setflag");
}

void Atom::set_mass(const char *file, int line, int itype, double value) {
    if (itype < 1 || itype > ntypes) error->all(file, line, "Invalid type for mass set");
    if (value <= 0.0) error->all(file, line, "Invalid mass value");
    mass[itype] = value;
    mass_setflag[itype] = 1;
}

void Error::all(const std::string &file, int line, const std::string &str) {
    fprintf(stderr, "ERROR: %s (%s:%d)\n", str.c_str(), file.c_str(), line);
    fflush(stderr);
    exit(3);
}

void Error::one(const std::string &file, int line, const std::string &str) {
    fprintf(stderr, "ERROR on proc 0: %s (%s:%d)\n", str.c_str(), file.c_str(), line);
    fflush(stderr);
    exit(3);
}

Neighbor::~Neighbor() {
    for (int i = 0; i < nrequest; ++i) delete requests[i];
}

int Neighbor::request(void *requestor, int instance) {
    if (nrequest == 8) error->all(FLERR, "Too many neighbor requests");
    NeighRequest *r = new NeighRequest;
    r->requestor = requestor;
    r->requestor_instance = instance;
    r->id = nrequest;
    requests[nrequest] = r;
    return nrequest++;
}

Pair::~Pair() {
    memory->destroy(eatom);
    memory->destroy(vatom);
}

void Pair::init_style() { neighbor->request(this, instance_me); }

// LAMMPS's Pair::ev_setup for these flag bits: zero the accumulators the
// flags select, size and zero eatom / vatom over owned atoms and (newton on)
// ghosts, and with vflag_global == VIRIAL_FDOTR hand the global virial to
// virial_fdotr_compute unless the style sets no_virial_fdotr_compute.
void Pair::ev_setup(int eflag, int vflag, int alloc) {
    evflag = 1;
    eflag_either = eflag;
    eflag_global = eflag & ENERGY_GLOBAL;
    eflag_atom = eflag & ENERGY_ATOM;
    vflag_global = vflag & (VIRIAL_PAIR | VIRIAL_FDOTR);
    vflag_atom = vflag & VIRIAL_ATOM;
    vflag_either = vflag_global || vflag_atom;
    evflag = eflag_either || vflag_either;
    if (eflag_atom && atom->nmax > maxeatom) {
        maxeatom = atom->nmax;
        if (alloc) {
            memory->destroy(eatom);
            memory->create(eatom, maxeatom, "pair:eatom");
        }
    }
    if (vflag_atom && atom->nmax > maxvatom) {
        maxvatom = atom->nmax;
        if (alloc) {
            memory->destroy(vatom);
            memory->create(vatom, maxvatom, 6, "pair:vatom");
        }
    }
    if (eflag_global) eng_vdwl = eng_coul = 0.0;
    if (vflag_global)
        for (int i = 0; i < 6; ++i) virial[i] = 0.0;
    const int n = atom->nlocal + (force->newton ? atom->nghost : 0);
    if (eflag_atom && alloc)
        for (int i = 0; i < n; ++i) eatom[i] = 0.0;
    if (vflag_atom && alloc)
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 6; ++k) vatom[i][k] = 0.0;
    if (vflag_global == VIRIAL_FDOTR && no_virial_fdotr_compute == 0) {
        vflag_fdotr = 1;
        vflag_global = 0;
        if (vflag_atom == 0) vflag_either = 0;
        if (vflag_either == 0 && eflag_either == 0) evflag = 0;
    } else {
        vflag_fdotr = 0;
    }
}

void Pair::ev_unset() {
    evflag = eflag_either = eflag_global = eflag_atom = 0;
    vflag_either = vflag_global = vflag_atom = vflag_fdotr = 0;
}

// global virial as sum over owned atoms and ghosts of x (x) f
void Pair::virial_fdotr_compute() {
    const int nall = atom->nlocal + atom->nghost;
    for (int i = 0; i < nall; ++i) {
        const double *x = atom->x[i], *f = atom->f[i];
        virial[0] += f[0] * x[0];
        virial[1] += f[1] * x[1];
        virial[2] += f[2] * x[2];
        virial[3] += f[1] * x[0];
        virial[4] += f[2] * x[0];
        virial[5] += f[2] * x[1];
    }
}

// A style that sets no flag and computes nothing: what LAMMPS's flags leave
// of the virial of a style that does not call virial_fdotr_compute.
class FdotrProbe : public Pair {
   public:
    explicit FdotrProbe(LAMMPS *lmp) : Pair(lmp) {}
    void compute(int eflag, int vflag) override { ev_init(eflag, vflag); }
    void settings(int, char **) override {}
    void coeff(int, char **) override {}
};

// ------------------------------------------------------------ driver
namespace {

template <typename T>
bool read_n(FILE *fp, T *dst, size_t n) {
    return fread(dst, sizeof(T), n, fp) == n;
}

template <typename T>
void write_n(FILE *fp, const T *src, size_t n) {
    if (fwrite(src, sizeof(T), n, fp) != n) {
        fprintf(stderr, "harness: short write\n");
        exit(2);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: %s <style> <potfile> <in.bin> <out.bin> <E1> [E2 ...]\n", argv[0]);
        return 2;
    }
    const std::string style = argv[1];
    FILE *in = fopen(argv[3], "rb");
    if (!in) {
        fprintf(stderr, "harness: cannot open %s\n", argv[3]);
        return 2;
    }
    int32_t hdr[8];
    if (!read_n(in, hdr, 8)) return 2;
    const int nlocal = hdr[0], nghost = hdr[1], ntypes = hdr[2], inum = hdr[3], nnbr = hdr[4];
    const int eflag = hdr[5], vflag = hdr[6], nsteps = hdr[7];
    const int nall = nlocal + nghost;
    if (argc != 5 + ntypes) {
        fprintf(stderr, "harness: %d element names for %d types\n", argc - 5, ntypes);
        return 2;
    }
    LAMMPS lmp;
    Atom *atom = lmp.atom;
    atom->grow(nall, ntypes);
    atom->nlocal = nlocal;
    atom->nghost = nghost;
    std::vector<int> ilist(inum), numneigh(nlocal > 0 ? nlocal : 1), nbr(nnbr > 0 ? nnbr : 1);
    std::vector<int32_t> tag32(nall);
    bool ok = read_n(in, &atom->x[0][0], 3 * static_cast<size_t>(nall)) && read_n(in, atom->type, nall) &&
              read_n(in, tag32.data(), nall) && read_n(in, ilist.data(), inum) &&
              read_n(in, numneigh.data(), nlocal) && read_n(in, nbr.data(), nnbr);
    fclose(in);
    if (!ok) {
        fprintf(stderr, "harness: short input\n");
        return 2;
    }
    for (int i = 0; i < nall; ++i) atom->tag[i] = static_cast<tagint>(tag32[i]);
    std::vector<int *> firstneigh(nlocal > 0 ? nlocal : 1, nullptr);
    size_t off = 0;
    for (int ii = 0; ii < inum; ++ii) {
        const int i = ilist[ii];
        firstneigh[i] = nbr.data() + off;
        off += numneigh[i];
    }
    if (off != static_cast<size_t>(nnbr)) {
        fprintf(stderr, "harness: neighbor counts do not add up\n");
        return 2;
    }
    NeighList list;
    list.inum = inum;
    list.ilist = ilist.data();
    list.numneigh = numneigh.data();
    list.firstneigh = firstneigh.data();

    Pair *pair = nullptr;
    if (style == "mlip_gtinv/hip" || style == "mlip_pair/hip")
        pair = new PairMLIPGtinvHIP(&lmp);
    else if (style == "fdotr_probe")
        pair = new FdotrProbe(&lmp);
    else {
        fprintf(stderr, "harness: unknown style %s\n", style.c_str());
        return 2;
    }
    // pair_style <style>; pair_coeff * * <potfile> <E1> ...
    pair->settings(0, nullptr);
    std::vector<std::string> sargs = {"*", "*", argv[2]};
    for (int t = 0; t < ntypes; ++t) sargs.push_back(argv[5 + t]);
    std::vector<char *> cargs;
    for (auto &a : sargs) cargs.push_back(&a[0]);
    if (style != "fdotr_probe") pair->coeff(static_cast<int>(cargs.size()), cargs.data());
    // init: list request, cutoffs
    pair->init_style();
    double cutoff = 0.0;
    if (style != "fdotr_probe")
        for (int i = 1; i <= ntypes; ++i)
            for (int j = i; j <= ntypes; ++j) cutoff = std::fmax(cutoff, pair->init_one(i, j));
    pair->init_list(0, &list);
    const NeighRequest *rq = lmp.neighbor->nrequest > 0 ? lmp.neighbor->requests[0] : nullptr;

    FILE *out = fopen(argv[4], "wb");
    if (!out) {
        fprintf(stderr, "harness: cannot open %s\n", argv[4]);
        return 2;
    }
    std::vector<double> eat(nall), vat(6 * static_cast<size_t>(nall));
    int32_t flags[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int step = 0; step < nsteps; ++step) {
        // Verlet::force_clear with newton on: owned atoms and ghosts
        for (int i = 0; i < nall; ++i) atom->f[i][0] = atom->f[i][1] = atom->f[i][2] = 0.0;
        pair->compute(eflag, vflag);
        if (step == 0) {
            flags[0] = rq ? rq->full : -1;
            flags[1] = rq ? rq->half : -1;
            flags[2] = lmp.force->newton_pair;
            flags[3] = pair->vflag_global;
            flags[4] = pair->vflag_fdotr;
            flags[5] = pair->eflag_atom;
            flags[6] = pair->vflag_atom;
            write_n(out, flags, 8);
            write_n(out, &cutoff, 1);
            write_n(out, atom->mass + 1, ntypes);
        }
        for (int i = 0; i < nall; ++i) {
            eat[i] = pair->eflag_atom ? pair->eatom[i] : 0.0;
            for (int k = 0; k < 6; ++k) vat[6 * i + k] = pair->vflag_atom ? pair->vatom[i][k] : 0.0;
        }
        write_n(out, &pair->eng_vdwl, 1);
        write_n(out, pair->virial, 6);
        write_n(out, &atom->f[0][0], 3 * static_cast<size_t>(nall));
        write_n(out, eat.data(), nall);
        write_n(out, vat.data(), 6 * static_cast<size_t>(nall));
    }
    fclose(out);
    delete pair;
    return 0;
}
