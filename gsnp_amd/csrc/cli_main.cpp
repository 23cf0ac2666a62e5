// gsnp -- SOAPsnp-compatible command line (cpp/main.cc:49-429) over libgsnp.
// Options follow the reference's getopt string; -q and -F are refused as in
// the reference.  -T (target regions), which the reference refuses
// (cpp/main.cc:295-306), calls only the windows the regions touch.  Extra:
// -B <windows per launch> selects the batched path.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unistd.h>

#include "../../include/gsnp_abi.h"

static int usage()
{
    std::fprintf(stderr,
                 "gsnp (MI355X)\n"
                 "Compulsory: -i <sorted SOAP alignments> -d <reference FASTA> -o <output>\n"
                 "Optional: -z <char q=0> -g <global dependency> -p <PCR dependency> -r <altHOM prior>\n"
                 "  -e <HET prior> -t (Ti/Tv 2:1) -s <known SNPs> -2 (refine with known SNPs)\n"
                 "  -a/-b/-j/-k <validated/unvalidated priors> -u (rank-sum test) -m (monoploid)\n"
                 "  -M <write matrix> -I <read matrix> -L <read length> -Q <char max quality>\n"
                 "  -T <target regions: chr start end> -C (compressed column output)\n"
                 "  -B <n> (batched path, n windows per launch)\n");
    return 1;
}

int main(int argc, char** argv)
{
    gsnp_params p;
    gsnp_params_default(&p);
    std::string in, fa, out, snp, min, mout, regions;
    int compress = 0, batched = 0, c;
    // -M and -I share one matrix stream in the reference (cpp/main.cc:236-260):
    // the later flag decides its direction, and any -I marks the matrix as read
    char matrix_last = 0;
    while ((c = getopt(argc, argv, "i:d:o:z:g:p:r:e:ts:2a:b:j:k:unmqM:I:L:Q:S:F:E:T:h:CB:")) != -1) {
        switch (c) {
        case 'i': in = optarg; break;
        case 'd': fa = optarg; break;
        case 'o': out = optarg; break;
        case 'z':
            p.q_min = optarg[0];
            p.q_max = p.q_min + 40;
            break;
        case 'g': p.global_dependency = std::log10(std::atof(optarg)); break;
        case 'p': p.pcr_dependency = std::log10(std::atof(optarg)); break;
        case 'r': p.althom_novel_r = std::atof(optarg); break;
        case 'e': p.het_novel_r = std::atof(optarg); break;
        case 't': p.transition_dominant = 1; break;
        case 's': snp = optarg; break;
        case '2': p.refine_mode = 1; break;
        case 'a': p.althom_val_r = std::atof(optarg); break;
        case 'b': p.het_val_r = std::atof(optarg); break;
        case 'j': p.althom_unval_r = std::atof(optarg); break;
        case 'k': p.het_unval_r = std::atof(optarg); break;
        case 'u': p.rank_sum_mode = 1; break;
        case 'n': break;
        case 'm': p.is_monoploid = 1; break;
        case 'q': std::printf("ERROR: The -q parameter currently is not supported.\n"); return 0;
        case 'M':
            mout = optarg;
            matrix_last = 'M';
            break;
        case 'I':
            min = optarg;
            matrix_last = 'I';
            break;
        case 'S': break;
        case 'L': p.read_length = std::atoi(optarg); break;
        case 'Q': p.q_max = optarg[0]; break;
        case 'F': std::printf("ERROR: The -F parameter currently is not supported.\n"); return 0;
        case 'E': std::printf("ERROR: The -E parameter currently is not supported.\n"); break;
        case 'T': regions = optarg; break;
        case 'C': compress = 1; break;
        case 'B': batched = std::atoi(optarg) < 1 ? 1 : std::atoi(optarg); break;
        default: return usage();
        }
    }
    if (in.empty() || fa.empty() || out.empty()) return usage();
    if (!min.empty() && !mout.empty()) {
        if (matrix_last == 'M') {
            // the reference would read the matrix from the stream it just opened
            // for writing: nothing, leaving every entry 1.0
            std::fprintf(stderr, "gsnp: -M after -I reads no matrix in SOAPsnp; give -I or -M, not both\n");
            return 255;
        }
        // -I after -M: the -M file is created (empty) and the matrix is read from -I
        if (FILE* f = std::fopen(mout.c_str(), "w")) std::fclose(f);
        mout.clear();
    }
    int rc = gsnp_device_init(0);
    gsnp_job* job = nullptr;
    if (!rc)
        rc = gsnp_job_open(fa.c_str(), in.c_str(), snp.empty() ? nullptr : snp.c_str(),
                           min.empty() ? nullptr : min.c_str(), &p, &job);
    if (!rc && !mout.empty()) rc = gsnp_job_write_matrix(job, mout.c_str());
    if (!rc && !regions.empty()) rc = gsnp_job_set_regions(job, regions.c_str());
    if (!rc) rc = gsnp_job_write(job, out.c_str(), compress, batched);
    gsnp_job_close(job);
    if (rc) {
        std::fprintf(stderr, "gsnp: error %d: %s\n", rc, gsnp_last_error());
        return 255;
    }
    return 0;
}
