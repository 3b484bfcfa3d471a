// Compiles against include/lynx_chefsi.hpp and links liblynx_chefsi.so; on a
// host without a GPU it exercises the set-up calls through a host-only
// context and checks that GPU work reports an error instead of aborting.
#include "lynx_chefsi.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char **argv)
{
    const char *psp = argc > 1 ? argv[1] : "C.pot";
    lynx_mi355x::Hamiltonian H(-1);
    const double a = 10.26;
    const double lv[9] = {a, 0, 0, 0, a, 0, 0, 0, a};
    H.setup(24, 24, 24, lv, true, 12);
    std::vector<double> frac = {0, 0, 0, .25, .25, .25, .5, .5, 0, .75, .75, .25,
                                .5, 0, .5, .75, .25, .75, 0, .5, .5, .25, .75, .75};
    H.set_atoms({psp}, std::vector<int>(8, 0), frac, true);
    int nimg = 0, natom = 0, nproj = 0;
    long ngpos = 0, nchi = 0;
    lynx_mi355x::check(lynx_get_nonlocal_sizes(H.ctx(), &nimg, &ngpos, &nchi, &natom, &nproj), H.ctx());
    std::vector<double> veff(H.Nd(), 0.0);
    bool threw = false;
    try {
        std::vector<double> psi(H.Nd()), y(H.Nd());
        H.apply(psi.data(), veff.data(), y.data(), 1);
    } catch (const std::exception &e) {
        threw = true;
    }
    std::printf("%d %ld %d %d %d\n", nimg, ngpos, natom, nproj, threw ? 1 : 0);
    return (natom == 8 && nproj == 64 && threw) ? 0 : 1;
}
