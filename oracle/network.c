/*
 * Loader for the network tables written by the network compiler
 * (3d-pdr_amd/pdrchem/netgen.py, write_tab).  The tables carry the reader
 * semantics of read_species.F90:91-190 and read_rates.F90:53-91 already
 * applied.  Test infrastructure (see pdr_oracle.h).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pdr_oracle.h"

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "pdr_oracle: out of memory\n"); abort(); }
  return p;
}

int pdr_net_load(const char *path, pdr_net *net) {
  memset(net, 0, sizeof(*net));
  net->xlambda_spline = 1;
  FILE *fh = fopen(path, "r");
  if (!fh) return -1;
  char line[16384];
  int nspec = -1, nreac = -1, ns = 0, nr = 0, nst = 0, cap = 0;
  while (fgets(line, sizeof line, fh)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    if (line[0] == 'N') {
      if (sscanf(line + 1, "%d %d", &nspec, &nreac) != 2) goto bad;
      net->nspec = nspec; net->nreac = nreac; net->neq = nspec - 1;
      net->charge = xcalloc(nspec, sizeof(int));
      net->abund0 = xcalloc(nspec, sizeof(double));
      net->mass = xcalloc(nspec, sizeof(double));
      net->rtype = xcalloc(nreac, sizeof(int));
      net->dup = xcalloc(nreac, sizeof(int));
      net->gflag = xcalloc(nreac, sizeof(int));
      net->fa = xcalloc(nreac, sizeof(int));
      net->fb = xcalloc(nreac, sizeof(int));
      net->nh = xcalloc(nreac, sizeof(int));
      net->alpha = xcalloc(nreac, sizeof(double));
      net->beta = xcalloc(nreac, sizeof(double));
      net->gamma = xcalloc(nreac, sizeof(double));
      net->rtmin = xcalloc(nreac, sizeof(double));
      net->rtmax = xcalloc(nreac, sizeof(double));
      net->st_ptr = xcalloc(nreac + 1, sizeof(int));
      cap = 8 * nreac;
      net->st_idx = xcalloc(cap, sizeof(int));
      net->st_coef = xcalloc(cap, sizeof(int));
    } else if (line[0] == 'F') {
      if (sscanf(line + 1, "%d", &net->xlambda_spline) != 1) goto bad;
    } else if (line[0] == 'S') {
      int i, q; double ab, m;
      if (nspec < 0 || sscanf(line + 1, "%d %d %lf %lf", &i, &q, &ab, &m) != 4) goto bad;
      if (i != ns || i >= nspec) goto bad;
      net->charge[i] = q; net->abund0[i] = ab; net->mass[i] = m;
      ns++;
    } else if (line[0] == 'R') {
      int i, t, d, g, a, b, h, k, off;
      double al, be, ga, tmin, tmax;
      char *p = line + 1;
      if (nreac < 0 || sscanf(p, "%d %d %d %d %lf %lf %lf %lf %lf %d %d %d %d%n", &i, &t, &d, &g,
                 &al, &be, &ga, &tmin, &tmax, &a, &b, &h, &k, &off) != 13) goto bad;
      if (i != nr || i >= nreac) goto bad;
      p += off;
      net->rtype[i] = t; net->dup[i] = d; net->gflag[i] = g;
      net->alpha[i] = al; net->beta[i] = be; net->gamma[i] = ga;
      net->rtmin[i] = tmin; net->rtmax[i] = tmax;
      net->fa[i] = a; net->fb[i] = b; net->nh[i] = h;
      net->st_ptr[i] = nst;
      for (int j = 0; j < k; j++) {
        int si, c, o2;
        if (sscanf(p, "%d %d%n", &si, &c, &o2) != 2) goto bad;
        p += o2;
        if (nst >= cap) {
          cap *= 2;
          net->st_idx = realloc(net->st_idx, cap * sizeof(int));
          net->st_coef = realloc(net->st_coef, cap * sizeof(int));
          if (!net->st_idx || !net->st_coef) abort();
        }
        net->st_idx[nst] = si; net->st_coef[nst] = c; nst++;
      }
      nr++;
      net->st_ptr[nr] = nst;
    } else if (line[0] == 'O') {
      int nb, nh, off;
      char *p = line + 1;
      if (nspec < 0 || sscanf(p, "%d %d%n", &nb, &nh, &off) != 2 || nb != nspec) goto bad;
      p += off;
      net->nb = nb; net->nhead = nh;
      net->border_order = xcalloc(nb, sizeof(int));
      for (int j = 0; j < nb; j++) {
        int v, o2;
        if (sscanf(p, "%d%n", &v, &o2) != 1 || v < 0 || v >= nb) goto bad;
        p += o2;
        net->border_order[j] = v;
      }
    } else if (line[0] == 'E') {
      break;
    }
  }
  fclose(fh);
  if (ns != nspec || nr != nreac) { pdr_net_free(net); return -2; }
  return 0;
bad:
  fclose(fh);
  pdr_net_free(net);
  return -3;
}

void pdr_net_free(pdr_net *net) {
  free(net->charge); free(net->abund0); free(net->mass);
  free(net->rtype); free(net->dup); free(net->gflag);
  free(net->fa); free(net->fb); free(net->nh);
  free(net->alpha); free(net->beta); free(net->gamma);
  free(net->rtmin); free(net->rtmax);
  free(net->st_ptr); free(net->st_idx); free(net->st_coef);
  free(net->border_order);
  memset(net, 0, sizeof(*net));
}

pdr_net *pdr_net_open(const char *path) {
  pdr_net *n = xcalloc(1, sizeof(pdr_net));
  if (pdr_net_load(path, n) != 0) { free(n); return NULL; }
  return n;
}

void pdr_net_close(pdr_net *net) {
  if (!net) return;
  pdr_net_free(net);
  free(net);
}

int pdr_net_nspec(const pdr_net *net) { return net->nspec; }
int pdr_net_nreac(const pdr_net *net) { return net->nreac; }
