/*
 * deck.c — .peec deck reader and peec_t lifetime (host side of the drop-in boundary).
 *
 * Reads the reference's "key = value" deck format (OpenPEEC src/input_data.c:38-602) into a
 * peec_t: the same keys, defaults (input_data.c:49-54), validations and error messages, so a
 * deck accepted by the reference is accepted here with an identical netlist.  peec_free()
 * releases exactly the members the reference frees (input_data.c:615-662) and, in addition,
 * drops the device-side mirror of this peec_t (peec_gpu_release).
 */
#include "peec.h"
#include "peec_internal.h"

#define MAXTOKEN 100

/* grow a malloc'ed array so that index num is valid */
#define GROW(ptr, num, cap, type)                                                  \
	do {                                                                           \
		if ((num) >= (cap)) {                                                      \
			(cap) = (cap) ? (2 * (cap)) : 16;                                      \
			(ptr) = (type *)realloc((ptr), (size_t)(cap) * sizeof(type));          \
		}                                                                          \
	} while (0)

int tokenize(char *str, const char *tokensep, char *token[], int maxtoken)
{
	if ((str == NULL) || (maxtoken == 0)) return 0;
	int count = 0;
	for (char *t = strtok(str, tokensep); (t != NULL) && (count < maxtoken); t = strtok(NULL, tokensep)) {
		token[count++] = t;
	}
	token[count] = NULL;
	return count;
}

/* non-negative integer node id (reference input_data.c:22-31) */
static int parse_node(const char *s, int *err)
{
	char *end = NULL;
	const long v = strtol(s, &end, 10);
	if ((end == s) || (*end != '\0') || (v < 0) || (v > 100000000)) {
		*err = 1;
		return 0;
	}
	return (int)v;
}

static void see_node(peec_t *p, int id)
{
	if (id > p->maxid) p->maxid = id;
}

static double len3(const double *v)
{
	return sqrt((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]));
}

static double dotp(const double *a, const double *b)
{
	return (a[0] * b[0]) + (a[1] * b[1]) + (a[2] * b[2]);
}

typedef struct {
	int res, cap, ind, mut, src, port, wire, node, plate, panel, diel;
} caps_t;

/* two-node element with a positive value: resistor / capacitor */
static int read_rc(peec_t *p, caps_t *c, char **tok, int ntok, int iscap)
{
	int err = (ntok < 5);
	if (!err) {
		rc_t **arr = iscap ? &p->cap : &p->res;
		int *num = iscap ? &p->ncap : &p->nres;
		int *cap = iscap ? &c->cap : &c->res;
		GROW(*arr, *num, *cap, rc_t);
		rc_t *e = &(*arr)[*num];
		e->n1 = parse_node(tok[2], &err);
		e->n2 = parse_node(tok[3], &err);
		e->val = atof(tok[4]);
		if (e->val <= 0) err = 1;
		see_node(p, e->n1);
		see_node(p, e->n2);
		if (!err) (*num)++;
	}
	return err;
}

/* quad panel sanity: planar and convex with cyclic vertices (reference input_data.c:358-398);
 * returns 0 ok, 1 invalid, 2 already reported */
static int check_quad(const panel_t *e)
{
	double nrm[3] = {0, 0, 0};
	for (int k = 0; k < 4; k++) {
		const double *a = &e->v[3 * k];
		const double *b = &e->v[3 * ((k + 1) % 4)];
		nrm[0] += (a[1] * b[2]) - (a[2] * b[1]);
		nrm[1] += (a[2] * b[0]) - (a[0] * b[2]);
		nrm[2] += (a[0] * b[1]) - (a[1] * b[0]);
	}
	const double nl = len3(nrm);
	if (nl <= 0) return 1;
	double size = 0;
	for (int k = 0; k < 4; k++) {
		const double *a = &e->v[3 * k];
		const double *b = &e->v[3 * ((k + 1) % 4)];
		const double *c = &e->v[3 * ((k + 2) % 4)];
		double u[3], w[3], x[3];
		for (int i = 0; i < 3; i++) {
			u[i] = b[i] - a[i];
			w[i] = c[i] - b[i];
		}
		const double ul = len3(u);
		if (ul > size) size = ul;
		double d0 = 0;
		for (int i = 0; i < 3; i++) d0 += (b[i] - e->v[i]) * nrm[i] / nl;
		if (fabs(d0) > 1e-9 * (size + fabs(d0))) {
			printf("%s\n", "*** quad : vertices must be coplanar");
			return 2;
		}
		x[0] = (u[1] * w[2]) - (u[2] * w[1]);
		x[1] = (u[2] * w[0]) - (u[0] * w[2]);
		x[2] = (u[0] * w[1]) - (u[1] * w[0]);
		if (dotp(x, nrm) <= 0) {
			printf("%s\n", "*** quad : must be convex (vertices in cyclic order)");
			return 2;
		}
	}
	return 0;
}

int input_data(FILE *fp, peec_t *p)
{
	char line[BUFSIZ], keep[BUFSIZ];
	char *tok[MAXTOKEN + 1];
	caps_t c;
	memset(&c, 0, sizeof(c));
	int nline = 0;

	memset(p, 0, sizeof(peec_t));
	p->nodetol = 1e-8;
	p->gmin = 0;
	p->refnode = -1;
	p->tranatt = 40;
	p->ctol = 1e-6;

	while (fgets(line, sizeof(line), fp) != NULL) {
		if (strlen(line) <= 1) continue;
		if (line[0] == '#') continue;
		if (strstr(line, "\r\n") != NULL) line[strlen(line) - 2] = '\0';
		else if ((strchr(line, '\r') != NULL) || (strchr(line, '\n') != NULL)) line[strlen(line) - 1] = '\0';
		if (!strncmp(line, "end", 3)) break;
		strcpy(keep, line);
		const int nt = tokenize(line, " \t", tok, MAXTOKEN);
		if ((nline > 0) && ((nt < 3) || strcmp(tok[1], "="))) continue;
		const char *key = tok[0];
		int err = 0;

		if (nline == 0) {
			if ((key == NULL) || strcmp(key, "OpenPEEC")) {
				printf("%s\n", "*** not OpenPEEC data");
				return 1;
			}
		}
		else if (!strcmp(key, "title")) {
			const char *s = strchr(keep, '=') + 1;
			while ((*s == ' ') || (*s == '\t')) s++;
			strcpy(p->title, s);
		}
		else if (!strcmp(key, "resistor") || !strcmp(key, "capacitor")) {
			if (read_rc(p, &c, tok, nt, !strcmp(key, "capacitor"))) {
				printf("*** invalid %s data\n", key);
				return 1;
			}
		}
		else if (!strcmp(key, "inductor")) {
			err = (nt < 6);
			if (!err) {
				GROW(p->ind, p->nind, c.ind, ind_t);
				ind_t *e = &p->ind[p->nind];
				strncpy(e->name, tok[2], NAMELEN - 1);
				e->name[NAMELEN - 1] = '\0';
				e->n1 = parse_node(tok[3], &err);
				e->n2 = parse_node(tok[4], &err);
				e->val = atof(tok[5]);
				if (e->val <= 0) err = 1;
				see_node(p, e->n1);
				see_node(p, e->n2);
				if (!err) p->nind++;
			}
			if (err) {
				printf("*** invalid %s data\n", "inductor");
				return 1;
			}
		}
		else if (!strcmp(key, "mutual")) {
			err = (nt < 5);
			if (!err) {
				GROW(p->mut, p->nmut, c.mut, mut_t);
				mut_t *e = &p->mut[p->nmut];
				strncpy(e->name1, tok[2], NAMELEN - 1);
				e->name1[NAMELEN - 1] = '\0';
				strncpy(e->name2, tok[3], NAMELEN - 1);
				e->name2[NAMELEN - 1] = '\0';
				e->k = atof(tok[4]);
				e->l1 = e->l2 = -1;
				if ((e->k <= 0) || (e->k > 1)) err = 1;
				if (!err) p->nmut++;
			}
			if (err) {
				printf("*** invalid %s data\n", "mutual");
				return 1;
			}
		}
		else if (!strcmp(key, "vsource") || !strcmp(key, "isource")) {
			err = (nt < 5);
			if (!err) {
				GROW(p->src, p->nsrc, c.src, src_t);
				src_t *e = &p->src[p->nsrc];
				e->isvsrc = !strcmp(key, "vsource");
				e->n1 = parse_node(tok[2], &err);
				e->n2 = parse_node(tok[3], &err);
				e->amp = atof(tok[4]);
				e->phase = (nt > 5) ? atof(tok[5]) : 0;
				e->ibr = -1;
				see_node(p, e->n1);
				see_node(p, e->n2);
				if (!err) p->nsrc++;
			}
			if (err) {
				printf("*** invalid %s data\n", key);
				return 1;
			}
		}
		else if (!strcmp(key, "node")) {
			err = (nt < 6);
			if (!err) {
				GROW(p->ncid, p->nnodexyz, c.node, int);
				p->ncxyz = (double *)realloc(p->ncxyz, (size_t)c.node * 3 * sizeof(double));
				p->ncid[p->nnodexyz] = parse_node(tok[2], &err);
				for (int k = 0; k < 3; k++) p->ncxyz[(3 * p->nnodexyz) + k] = atof(tok[3 + k]);
				see_node(p, p->ncid[p->nnodexyz]);
				if (!err) p->nnodexyz++;
			}
			if (err) {
				printf("*** invalid %s data\n", "node");
				return 1;
			}
		}
		else if (!strcmp(key, "wire") || !strcmp(key, "bar")) {
			const int isbar = !strcmp(key, "bar");
			err = (nt < (isbar ? 12 : 11));
			if (!err) {
				GROW(p->wire, p->nwire, c.wire, wire_t);
				wire_t *e = &p->wire[p->nwire];
				memset(e, 0, sizeof(wire_t));
				for (int k = 0; k < 3; k++) {
					e->x1[k] = atof(tok[2 + k]);
					e->x2[k] = atof(tok[5 + k]);
				}
				if (isbar) {
					e->shape = SHAPE_BAR;
					e->width = atof(tok[8]);
					e->thick = atof(tok[9]);
					e->sigma = atof(tok[10]);
					e->ndiv = atoi(tok[11]);
					if ((e->width <= 0) || (e->thick <= 0)) err = 1;
				}
				else {
					e->shape = SHAPE_ROUND;
					e->radius = atof(tok[8]);
					e->sigma = atof(tok[9]);
					e->ndiv = atoi(tok[10]);
					if (e->radius <= 0) err = 1;
				}
				double d[3] = {e->x2[0] - e->x1[0], e->x2[1] - e->x1[1], e->x2[2] - e->x1[2]};
				if ((e->sigma < 0) || (e->ndiv < 1) || (len3(d) <= 0)) err = 1;
				if (!err) p->nwire++;
			}
			if (err) {
				printf("*** invalid %s data\n", key);
				return 1;
			}
		}
		else if (!strcmp(key, "plate")) {
			err = (nt < 15);
			if (!err) {
				GROW(p->plate, p->nplate, c.plate, plate_t);
				plate_t *e = &p->plate[p->nplate];
				memset(e, 0, sizeof(plate_t));
				for (int k = 0; k < 3; k++) {
					e->org[k] = atof(tok[2 + k]);
					e->ea[k] = atof(tok[5 + k]);
					e->eb[k] = atof(tok[8 + k]);
				}
				e->thick = atof(tok[11]);
				e->sigma = atof(tok[12]);
				e->ndiva = atoi(tok[13]);
				e->ndivb = atoi(tok[14]);
				e->ndivt = (nt >= 16) ? atoi(tok[15]) : 1;
				const double la = len3(e->ea), lb = len3(e->eb), ab = dotp(e->ea, e->eb);
				if ((la <= 0) || (lb <= 0) || (e->thick <= 0) || (e->sigma < 0) ||
				    (e->ndiva < 1) || (e->ndivb < 1) || (e->ndivt < 1)) err = 1;
				if (!err && (fabs(ab) > 1e-9 * la * lb)) {
					printf("%s\n", "*** plate : the two edge vectors must be perpendicular");
					return 1;
				}
				if (!err) p->nplate++;
			}
			if (err) {
				printf("*** invalid %s data\n", "plate");
				return 1;
			}
		}
		else if (!strcmp(key, "dielectric")) {
			err = (nt < 16);
			if (!err) {
				GROW(p->diel, p->ndiel, c.diel, diel_t);
				diel_t *e = &p->diel[p->ndiel];
				memset(e, 0, sizeof(diel_t));
				for (int k = 0; k < 3; k++) {
					e->org[k] = atof(tok[2 + k]);
					e->ea[k] = atof(tok[5 + k]);
					e->eb[k] = atof(tok[8 + k]);
				}
				e->thick = atof(tok[11]);
				e->epsr = atof(tok[12]);
				e->ndiva = atoi(tok[13]);
				e->ndivb = atoi(tok[14]);
				e->ndivt = atoi(tok[15]);
				e->tand = (nt >= 17) ? atof(tok[16]) : 0;
				if (e->tand < 0) err = 1;
				if (nt >= 19) {
					e->epsinf = atof(tok[17]);
					e->frelax = atof(tok[18]);
					if ((e->epsinf < 1) || (e->epsinf > e->epsr) || (e->frelax <= 0) || (e->tand != 0)) err = 1;
				}
				const double la = len3(e->ea), lb = len3(e->eb), ab = dotp(e->ea, e->eb);
				if ((la <= 0) || (lb <= 0) || (e->thick <= 0) || (e->epsr < 1) ||
				    (e->ndiva < 1) || (e->ndivb < 1) || (e->ndivt < 1)) err = 1;
				if (!err && (fabs(ab) > 1e-9 * la * lb)) {
					printf("%s\n", "*** dielectric : the two edge vectors must be perpendicular");
					return 1;
				}
				/* epsr == 1 is vacuum: no cells at all (reference input_data.c:331-332) */
				if (!err && (e->epsr > 1)) p->ndiel++;
			}
			if (err) {
				printf("*** invalid %s data\n", "dielectric");
				return 1;
			}
		}
		else if (!strcmp(key, "quad")) {
			err = (nt < 18);
			if (!err) {
				GROW(p->panel, p->npanel, c.panel, panel_t);
				panel_t *e = &p->panel[p->npanel];
				memset(e, 0, sizeof(panel_t));
				e->kind = PANEL_QUAD;
				for (int k = 0; k < 12; k++) e->v[k] = atof(tok[2 + k]);
				e->thick = atof(tok[14]);
				e->sigma = atof(tok[15]);
				e->ndiva = atoi(tok[16]);
				e->ndivb = atoi(tok[17]);
				if ((e->thick <= 0) || (e->sigma < 0) || (e->ndiva < 1) || (e->ndivb < 1)) err = 1;
				if (!err) {
					const int q = check_quad(e);
					if (q == 2) return 1;
					err = q;
				}
				if (!err) p->npanel++;
			}
			if (err) {
				printf("*** invalid %s data\n", "quad");
				return 1;
			}
		}
		else if (!strcmp(key, "disk")) {
			err = (nt < 13);
			if (!err) {
				GROW(p->panel, p->npanel, c.panel, panel_t);
				panel_t *e = &p->panel[p->npanel];
				memset(e, 0, sizeof(panel_t));
				e->kind = PANEL_DISK;
				for (int k = 0; k < 3; k++) {
					e->org[k] = atof(tok[2 + k]);
					e->nrm[k] = atof(tok[5 + k]);
				}
				e->radius = atof(tok[8]);
				e->thick = atof(tok[9]);
				e->sigma = atof(tok[10]);
				e->ndiva = atoi(tok[11]);
				e->ndivb = atoi(tok[12]);
				/* the centre dual ring has 2 nsec vertices, so nsec <= POLY_MAX / 2 */
				if ((len3(e->nrm) <= 0) || (e->radius <= 0) || (e->thick <= 0) || (e->sigma < 0) ||
				    (e->ndiva < 1) || (e->ndivb < 3) || (e->ndivb > POLY_MAX / 2)) err = 1;
				if (!err) p->npanel++;
			}
			if (err) {
				printf("*** invalid %s data\n", "disk");
				return 1;
			}
		}
		else if (!strcmp(key, "port")) {
			err = (nt < 5);
			if (!err) {
				GROW(p->port, p->nport, c.port, port_t);
				port_t *e = &p->port[p->nport];
				e->n1 = parse_node(tok[2], &err);
				e->n2 = parse_node(tok[3], &err);
				e->z0 = atof(tok[4]);
				if (e->n1 == e->n2) err = 1;
				see_node(p, e->n1);
				see_node(p, e->n2);
				if (!err) p->nport++;
			}
			if (err) {
				printf("*** invalid %s data\n", "port");
				return 1;
			}
		}
		else if (!strcmp(key, "frequency")) {
			err = (nt < 5);
			if (!err) {
				p->f0 = atof(tok[2]);
				p->f1 = atof(tok[3]);
				const int ndiv = atoi(tok[4]);
				p->nfreq = ndiv + 1;
				if ((nt > 5) && !strcmp(tok[5], "log")) p->flog = 1;
				if ((p->f0 <= 0) || (p->f1 < p->f0) || (ndiv < 0)) err = 1;
			}
			if (err) {
				printf("*** invalid %s data\n", "frequency");
				return 1;
			}
		}
		else if (!strcmp(key, "nodetol")) {
			p->nodetol = atof(tok[2]);
			if (p->nodetol <= 0) {
				printf("*** invalid %s data\n", "nodetol");
				return 1;
			}
		}
		else if (!strcmp(key, "skineffect"))   p->skin = atoi(tok[2]);
		else if (!strcmp(key, "capacitance"))  p->capacitance = atoi(tok[2]);
		else if (!strcmp(key, "distribution")) p->dist = atoi(tok[2]);
		else if (!strcmp(key, "retardation"))  p->retardation = atoi(tok[2]);
		else if (!strcmp(key, "acceleration")) p->accel = atoi(tok[2]);
		else if (!strcmp(key, "grading"))      p->grading = atoi(tok[2]);
		else if (!strcmp(key, "transient")) {
			p->tran = atoi(tok[2]);
			if (nt >= 4) p->tranatt = atof(tok[3]);
			if (p->tranatt <= 0) {
				printf("*** invalid %s data\n", "transient");
				return 1;
			}
		}
		else if (!strcmp(key, "planewave")) {
			err = (nt < 6);
			if (!err) {
				p->pwth = atof(tok[2]);
				p->pwph = atof(tok[3]);
				p->pwpol = atoi(tok[4]);
				p->pwamp = atof(tok[5]);
				p->pwphase = (nt >= 7) ? atof(tok[6]) : 0;
				if (((p->pwpol != 1) && (p->pwpol != 2)) || (p->pwamp <= 0)) err = 1;
				if (!err) p->pw = 1;
			}
			if (err) {
				printf("*** invalid %s data\n", "planewave");
				return 1;
			}
		}
		else if (!strcmp(key, "farfield")) {
			err = (nt < 4);
			if (!err) {
				p->ffnth = atoi(tok[2]);
				p->ffnph = atoi(tok[3]);
				if ((p->ffnth < 2) || (p->ffnph < 2)) err = 1;
			}
			if (err) {
				printf("*** invalid %s data\n", "farfield");
				return 1;
			}
		}
		else if (!strcmp(key, "compression")) {
			p->compress = atoi(tok[2]);
			if (nt >= 4) p->ctol = atof(tok[3]);
			if (p->ctol <= 0) {
				printf("*** invalid %s data\n", "compression");
				return 1;
			}
		}
		else if (!strcmp(key, "groundplane")) {
			p->gp = 1;
			p->gpz = atof(tok[2]);
		}
		else if (!strcmp(key, "gmin")) {
			p->gmin = atof(tok[2]);
			if (p->gmin < 0) {
				printf("*** invalid %s data\n", "gmin");
				return 1;
			}
		}
		/* unknown keys are ignored (forward compatibility) */
		nline++;
	}

	for (int i = 0; i < p->nmut; i++) {
		mut_t *e = &p->mut[i];
		for (int j = 0; j < p->nind; j++) {
			if (!strcmp(e->name1, p->ind[j].name)) e->l1 = j;
			if (!strcmp(e->name2, p->ind[j].name)) e->l2 = j;
		}
		if ((e->l1 < 0) || (e->l2 < 0) || (e->l1 == e->l2)) {
			printf("*** mutual : unknown or same inductor : %s %s\n", e->name1, e->name2);
			return 1;
		}
	}
	if (p->nport <= 0) {
		printf("%s\n", "*** no port data");
		return 1;
	}
	if (p->nfreq <= 0) {
		printf("%s\n", "*** no frequency data");
		return 1;
	}
	if ((p->nres + p->ncap + p->nind + p->nwire + p->nplate + p->npanel) <= 0) {
		printf("%s\n", "*** no element data (resistor/capacitor/inductor/wire/bar/plate/polygon/disk)");
		return 1;
	}
	if ((p->ndiel > 0) && !p->capacitance) {
		printf("%s\n", "*** dielectric requires capacitance = 1");
		return 1;
	}
	if (p->compress && p->accel) {
		printf("%s\n", "*** compression = 1 and acceleration = 1 are mutually exclusive");
		return 1;
	}
	if (p->compress && ((p->nwire + p->nplate + p->npanel) <= 0)) {
		printf("%s\n", "*** compression = 1 requires conductor geometry (wire/bar/plate/quad/disk)");
		return 1;
	}
	return 0;
}

void peec_free(peec_t *p)
{
	if (p == NULL) return;
	peec_gpu_release(p);

	free(p->plate);
	free(p->panel);
	free(p->diel);
	free(p->res);
	free(p->cap);
	free(p->ind);
	free(p->mut);
	free(p->src);
	free(p->port);
	free(p->wire);
	free(p->ncid);
	free(p->ncxyz);

	free(p->seg);
	free(p->chg);
	free(p->chgnode);
	free(p->gid);
	free(p->gxyz);

	free(p->lp);
	free(p->cellid);
	free(p->cellof);
	free(p->carea);
	free(p->csoff);
	free(p->csidx);
	free(p->cpt);
	free(p->cmat);

	free(p->nodemap);
	free(p->zin);
	free(p->zmat);
	free(p->smat);
	free(p->segi);
	free(p->cellq);
	free(p->voc);
	free(p->segipw);

	memset(p, 0, sizeof(peec_t));
}
