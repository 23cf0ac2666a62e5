/*
 * oracle_cli.c -- command line for the CPU oracle (test infrastructure only).
 * Flags follow cpp/main.cc:104 (SOAPsnp's getopt string) for the options the
 * calling path uses; extra flags: -D <prefix> decompress, -X <n> stop after n
 * sites (bounded timing sample), -R print the call-stage seconds, -J <n> call
 * the windows of text output on n threads.  -T <regions> calls only the
 * windows the regions touch.
 */
#include "gsnp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

int main(int argc, char** argv)
{
    orc_params p;
    orc_params_default(&p);
    const char *in = NULL, *fa = NULL, *out = NULL, *snp = NULL, *min = NULL, *mout = NULL, *dec = NULL,
               *regions = NULL;
    int compress = 0, report = 0, threads = 1;
    long long max_sites = 0;
    int c;
    while ((c = getopt(argc, argv, "i:d:o:z:g:p:r:e:ts:2a:b:j:k:umM:I:L:Q:CD:X:RJ:T:")) != -1) {
        switch (c) {
        case 'i': in = optarg; break;
        case 'd': fa = optarg; break;
        case 'o': out = optarg; break;
        case 'z': p.q_min = optarg[0]; p.q_max = p.q_min + 40; break;
        case 'g': p.global_dependency = log10(atof(optarg)); break;
        case 'p': p.pcr_dependency = log10(atof(optarg)); break;
        case 'r': p.althom_novel_r = atof(optarg); break;
        case 'e': p.het_novel_r = atof(optarg); break;
        case 't': p.transition_dominant = 1; break;
        case 's': snp = optarg; break;
        case '2': p.refine_mode = 1; break;
        case 'a': p.althom_val_r = atof(optarg); break;
        case 'b': p.het_val_r = atof(optarg); break;
        case 'j': p.althom_unval_r = atof(optarg); break;
        case 'k': p.het_unval_r = atof(optarg); break;
        case 'u': p.rank_sum_mode = 1; break;
        case 'm': p.is_monoploid = 1; break;
        case 'M': mout = optarg; break;
        case 'I': min = optarg; break;
        case 'L': p.read_length = atoi(optarg); break;
        case 'Q': p.q_max = optarg[0]; break;
        case 'C': compress = 1; break;
        case 'D': dec = optarg; break;
        case 'X': max_sites = atoll(optarg); break;
        case 'R': report = 1; break;
        case 'J': threads = atoi(optarg) < 1 ? 1 : atoi(optarg); break;
        case 'T': regions = optarg; break;
        default: fprintf(stderr, "unknown option\n"); return 1;
        }
    }
    if (dec) {
        if (!out) return 1;
        int rc = orc_decompress(dec, out);
        if (rc) fprintf(stderr, "error %d: %s\n", rc, orc_last_error());
        return rc;
    }
    if (!in || !fa || !out) {
        fprintf(stderr, "usage: gsnp_oracle -i soap -d fasta -o out [-s dbsnp] [-C] ...\n");
        return 1;
    }
    orc_run_stats st;
    int rc = orc_run_opts(fa, in, snp, min, mout, &p, out, compress, max_sites, threads, regions, &st);
    if (rc) {
        fprintf(stderr, "error %d: %s\n", rc, orc_last_error());
        return rc;
    }
    if (report) printf("{\"sites\": %llu, \"windows\": %llu, \"seconds_call\": %.6f}\n", (unsigned long long)st.sites,
                       (unsigned long long)st.windows, st.seconds_call);
    return 0;
}
