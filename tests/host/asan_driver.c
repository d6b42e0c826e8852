/*
 * The host-side paths of libfmd.so that run without a GPU, for a build of the
 * host sources with AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_host_sanitizers.py): the setfl loader on every table named on
 * the command line, LJ and Morse pairs, the lattice builders, Maxwell
 * velocities, translation, configuration and state files, and the state
 * reader.  Usage: asan_driver <save dir ending in '/'> <setfl> [<setfl> ...]
 */
#include <stdio.h>
#include <string.h>
#include "fmd.h"

static fmd_t *make(void)
{
    fmd_t *md = fmd_create();
    fmd_io_setShowErrorMessages(md, true);
    fmd_box_setSize(md, 40.0, 40.0, 40.0);
    fmd_box_setPBC(md, true, true, true);
    fmd_box_setSubdomains(md, 1, 1, 1);
    char cu[] = "Cu", ar[] = "Ar";
    fmd_string_t names[] = {cu, ar};
    fmd_real_t masses[] = {63.546, 39.948};
    fmd_matt_setAtomKinds(md, 2, names, masses);
    return md;
}

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    char *dir = argv[1];
    fmd_t *md = make();
    for (int i = 2; i < argc; i++)
    {
        fmd_pot_t *p = fmd_pot_eam_alloy_load(md, argv[i]);
        if (!p)
        {
            fprintf(stderr, "setfl load failed: %s\n", argv[i]);
            return 3;
        }
        printf("%s: cutoff %.6f\n", argv[i], fmd_pot_eam_getCutoffRadius(md, p));
        if (i == 2) fmd_pot_apply(md, 0, 0, p);
    }
    fmd_pot_lj_apply(md, 1, 1, 3.405, 0.0104, 2.5 * 3.405);
    fmd_pot_morse_apply(md, 0, 1, 0.3429, 1.3588, 2.866, 7.0);
    fmd_matt_makeCuboidFCC(md, 0.0, 0.0, 0.0, 6, 6, 6, 3.615, 0, 0, 300.0);
    fmd_matt_makeCuboidFCC(md, 22.0, 0.0, 0.0, 3, 6, 6, 5.26, 1, 1, 100.0);
    fmd_matt_giveMaxwellDistribution(md, 1, 50.0);
    fmd_matt_translate(md, 1, 0.5, -0.25, 0.125);
    fmd_rtriple_t lo, hi;
    fmd_matt_findLimits(md, -1, lo, hi);
    fmd_io_setSaveDirectory(md, dir);
    fmd_io_setSaveConfigMode(md, FMD_SCM_XYZ_ATOMSNUM);
    fmd_matt_saveConfiguration(md);
    fmd_io_setSaveConfigMode(md, FMD_SCM_CSV);
    fmd_matt_saveConfiguration(md);
    fmd_io_setSaveConfigMode(md, FMD_SCM_VTF);
    fmd_matt_saveConfiguration(md);
    fmd_io_saveState(md, "asan.stt");
    fmd_free(md);

    char path[4096];
    snprintf(path, sizeof path, "%sasan.stt", dir);
    md = make();   /* the reader bins into the grid, which needs the cutoff (misc.c:41-62) */
    fmd_pot_lj_apply(md, 0, 0, 2.3, 0.4, 5.5);
    fmd_pot_lj_apply(md, 1, 1, 3.405, 0.0104, 2.5 * 3.405);
    fmd_pot_morse_apply(md, 0, 1, 0.3429, 1.3588, 2.866, 7.0);
    fmd_io_loadState(md, path, true);
    fmd_io_setSaveDirectory(md, dir);
    fmd_io_saveState(md, "asan2.stt");
    fmd_free(md);
    printf("host paths ok\n");
    return 0;
}
