/*
 * deck.c — reader of .peec decks and peec_t lifetime (host side of the drop-in boundary).
 *
 * The deck grammar, defaults, validations and messages are the reference's interface
 * (OpenPEEC src/input_data.c:38-602): a deck the reference accepts gives the same netlist here,
 * and a deck it rejects is rejected with the same "*** ..." line.  The reader is organised
 * around a keyword table: every line "key = fields..." is dispatched through KEYWORDS to a
 * handler that pulls typed fields from a cursor (each missing or malformed field marks the line
 * bad); flag keywords and scalar limits are table entries with a member offset, not code.
 * Skin effect (skineffect = 1) is not part of this path and is refused.
 * peec_free() releases what the reference frees (input_data.c:615-662) plus the device mirror.
 */
#include <stddef.h>
#include "peec.h"
#include "peec_internal.h"

#define MAX_FIELDS 100

/* tokenize (utils.c:10): split str in place at any of the separator characters; at most
 * maxtoken tokens, token[count] = NULL; returns the count */
int tokenize(char *str, const char *tokensep, char *token[], int maxtoken)
{
	if ((str == NULL) || (maxtoken == 0)) return 0;
	int count = 0;
	char *s = str;
	while (count < maxtoken) {
		s += strspn(s, tokensep);
		if (*s == '\0') break;
		token[count++] = s;
		s += strcspn(s, tokensep);
		if (*s == '\0') break;
		*s++ = '\0';
	}
	token[count] = NULL;
	return count;
}

/* ------------------------------------------------------------------ typed field cursor --- */

typedef struct {
	char **f;       /* the fields after "key =" */
	int n, next;
	int bad;        /* a field was missing, malformed or out of range */
} cursor_t;

static int more(const cursor_t *c, int k)
{
	return c->n - c->next >= k;
}

static double real_field(cursor_t *c)
{
	if (!more(c, 1)) {
		c->bad = 1;
		return 0;
	}
	return atof(c->f[c->next++]);
}

static int int_field(cursor_t *c)
{
	if (!more(c, 1)) {
		c->bad = 1;
		return 0;
	}
	return atoi(c->f[c->next++]);
}

/* a node id: a whole non-negative decimal number no larger than 1e8 */
static int node_field(cursor_t *c, peec_t *p)
{
	if (!more(c, 1)) {
		c->bad = 1;
		return 0;
	}
	const char *s = c->f[c->next++];
	char *end = NULL;
	const long v = strtol(s, &end, 10);
	if ((end == s) || (*end != '\0') || (v < 0) || (v > 100000000)) {
		c->bad = 1;
		return 0;
	}
	if (v > p->maxid) p->maxid = (int)v;
	return (int)v;
}

static void name_field(cursor_t *c, char *dst)
{
	if (!more(c, 1)) {
		c->bad = 1;
		return;
	}
	snprintf(dst, NAMELEN, "%s", c->f[c->next++]);
}

static void vec_field(cursor_t *c, double *v, int k)
{
	for (int i = 0; i < k; i++) v[i] = real_field(c);
}

static void require(cursor_t *c, int ok)
{
	if (!ok) c->bad = 1;
}

/* --------------------------------------------------------------------- helpers ----------- */

static double norm3(const double *v)
{
	return sqrt((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]));
}

static double inner3(const double *a, const double *b)
{
	return (a[0] * b[0]) + (a[1] * b[1]) + (a[2] * b[2]);
}

static void cross3(const double *a, const double *b, double *x)
{
	x[0] = (a[1] * b[2]) - (a[2] * b[1]);
	x[1] = (a[2] * b[0]) - (a[0] * b[2]);
	x[2] = (a[0] * b[1]) - (a[1] * b[0]);
}

/* room for one more element in a growing array; returns the new (zeroed) slot */
static void *append(void **arr, int count, int *cap, size_t elem)
{
	if (count >= *cap) {
		const int nc = (*cap > 0) ? 2 * *cap : 16;
		void *q = realloc(*arr, (size_t)nc * elem);
		if (q == NULL) return NULL;
		*arr = q;
		*cap = nc;
	}
	void *slot = (char *)*arr + (size_t)count * elem;
	memset(slot, 0, elem);
	return slot;
}

/* capacities of the growing arrays while a deck is read */
typedef struct {
	int res, cap, ind, mut, src, port, wire, node, plate, panel, diel;
	const char *line;   /* the current line as read (for title) */
} deck_t;

enum { KW_OK = 0, KW_BAD = 1, KW_REPORTED = 2 };

typedef int (*kw_fn)(peec_t *p, deck_t *d, cursor_t *c, size_t arg);

/* --------------------------------------------------------------------- handlers ---------- */

static int kw_title(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)c;
	(void)arg;
	const char *s = strchr(d->line, '=') + 1;
	s += strspn(s, " \t");
	snprintf(p->title, sizeof(p->title), "%s", s);
	return KW_OK;
}

/* resistor / capacitor: n1 n2 value > 0 (arg 1: capacitor) */
static int kw_lumped(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	rc_t **arr = arg ? &p->cap : &p->res;
	int *num = arg ? &p->ncap : &p->nres;
	rc_t *e = (rc_t *)append((void **)arr, *num, arg ? &d->cap : &d->res, sizeof(rc_t));
	if (e == NULL) return KW_BAD;
	e->n1 = node_field(c, p);
	e->n2 = node_field(c, p);
	e->val = real_field(c);
	require(c, e->val > 0);
	if (c->bad) return KW_BAD;
	(*num)++;
	return KW_OK;
}

static int kw_inductor(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)arg;
	ind_t *e = (ind_t *)append((void **)&p->ind, p->nind, &d->ind, sizeof(ind_t));
	if (e == NULL) return KW_BAD;
	name_field(c, e->name);
	e->n1 = node_field(c, p);
	e->n2 = node_field(c, p);
	e->val = real_field(c);
	require(c, e->val > 0);
	if (c->bad) return KW_BAD;
	p->nind++;
	return KW_OK;
}

/* mutual = L1name L2name k, 0 < k <= 1 (inductor names resolved after the deck) */
static int kw_mutual(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)arg;
	mut_t *e = (mut_t *)append((void **)&p->mut, p->nmut, &d->mut, sizeof(mut_t));
	if (e == NULL) return KW_BAD;
	name_field(c, e->name1);
	name_field(c, e->name2);
	e->k = real_field(c);
	e->l1 = e->l2 = -1;
	require(c, (e->k > 0) && (e->k <= 1));
	if (c->bad) return KW_BAD;
	p->nmut++;
	return KW_OK;
}

/* vsource / isource = n1 n2 amplitude [phase deg] (arg 1: voltage source) */
static int kw_source(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	src_t *e = (src_t *)append((void **)&p->src, p->nsrc, &d->src, sizeof(src_t));
	if (e == NULL) return KW_BAD;
	e->isvsrc = (int)arg;
	e->n1 = node_field(c, p);
	e->n2 = node_field(c, p);
	e->amp = real_field(c);
	e->phase = more(c, 1) ? real_field(c) : 0;
	e->ibr = -1;
	if (c->bad) return KW_BAD;
	p->nsrc++;
	return KW_OK;
}

/* node = id x y z: a node placed at a point (geometry snaps to it) */
static int kw_node(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)arg;
	int *id = (int *)append((void **)&p->ncid, p->nnodexyz, &d->node, sizeof(int));
	double *q = (double *)realloc(p->ncxyz, (size_t)d->node * 3 * sizeof(double));
	if ((id == NULL) || (q == NULL)) return KW_BAD;
	p->ncxyz = q;
	*id = node_field(c, p);
	vec_field(c, &p->ncxyz[3 * (size_t)p->nnodexyz], 3);
	if (c->bad) return KW_BAD;
	p->nnodexyz++;
	return KW_OK;
}

/* wire = x1 x2 radius sigma ndiv; bar = x1 x2 width thick sigma ndiv (arg: shape) */
static int kw_wire(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	wire_t *e = (wire_t *)append((void **)&p->wire, p->nwire, &d->wire, sizeof(wire_t));
	if (e == NULL) return KW_BAD;
	vec_field(c, e->x1, 3);
	vec_field(c, e->x2, 3);
	e->shape = (int)arg;
	if (arg == SHAPE_BAR) {
		e->width = real_field(c);
		e->thick = real_field(c);
		require(c, (e->width > 0) && (e->thick > 0));
	}
	else {
		e->radius = real_field(c);
		require(c, e->radius > 0);
	}
	e->sigma = real_field(c);
	e->ndiv = int_field(c);
	const double d12[3] = {e->x2[0] - e->x1[0], e->x2[1] - e->x1[1], e->x2[2] - e->x1[2]};
	require(c, (e->sigma >= 0) && (e->ndiv >= 1) && (norm3(d12) > 0));
	if (c->bad) return KW_BAD;
	p->nwire++;
	return KW_OK;
}

/* org, edge a, edge b of a rectangle, perpendicular edges; returns KW_REPORTED otherwise */
static int rect_frame(cursor_t *c, double *org, double *ea, double *eb, const char *what)
{
	vec_field(c, org, 3);
	vec_field(c, ea, 3);
	vec_field(c, eb, 3);
	const double la = norm3(ea), lb = norm3(eb);
	require(c, (la > 0) && (lb > 0));
	if (!c->bad && (fabs(inner3(ea, eb)) > 1e-9 * la * lb)) {
		printf("*** %s : the two edge vectors must be perpendicular\n", what);
		return KW_REPORTED;
	}
	return KW_OK;
}

/* plate = org ea eb thick sigma na nb [nt] */
static int kw_plate(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)arg;
	plate_t *e = (plate_t *)append((void **)&p->plate, p->nplate, &d->plate, sizeof(plate_t));
	if (e == NULL) return KW_BAD;
	double org[3], ea[3], eb[3];
	vec_field(c, org, 3);
	vec_field(c, ea, 3);
	vec_field(c, eb, 3);
	memcpy(e->org, org, sizeof(org));
	memcpy(e->ea, ea, sizeof(ea));
	memcpy(e->eb, eb, sizeof(eb));
	e->thick = real_field(c);
	e->sigma = real_field(c);
	e->ndiva = int_field(c);
	e->ndivb = int_field(c);
	e->ndivt = more(c, 1) ? int_field(c) : 1;
	const double la = norm3(ea), lb = norm3(eb);
	require(c, (la > 0) && (lb > 0) && (e->thick > 0) && (e->sigma >= 0) && (e->ndiva >= 1) && (e->ndivb >= 1) && (e->ndivt >= 1));
	if (c->bad) return KW_BAD;
	if (fabs(inner3(ea, eb)) > 1e-9 * la * lb) {
		printf("%s\n", "*** plate : the two edge vectors must be perpendicular");
		return KW_REPORTED;
	}
	p->nplate++;
	return KW_OK;
}

/* dielectric = org ea eb thick epsr na nb nt [tand [epsinf frelax]]; epsr = 1 is vacuum and
 * adds nothing */
static int kw_dielectric(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)arg;
	diel_t *e = (diel_t *)append((void **)&p->diel, p->ndiel, &d->diel, sizeof(diel_t));
	if (e == NULL) return KW_BAD;
	vec_field(c, e->org, 3);
	vec_field(c, e->ea, 3);
	vec_field(c, e->eb, 3);
	e->thick = real_field(c);
	e->epsr = real_field(c);
	e->ndiva = int_field(c);
	e->ndivb = int_field(c);
	e->ndivt = int_field(c);
	e->tand = more(c, 1) ? real_field(c) : 0;
	require(c, e->tand >= 0);
	if (more(c, 2)) {   /* Debye relaxation: eps_inf in [1, eps_r], f_relax > 0, no tan delta */
		e->epsinf = real_field(c);
		e->frelax = real_field(c);
		require(c, (e->epsinf >= 1) && (e->epsinf <= e->epsr) && (e->frelax > 0) && (e->tand == 0));
	}
	const double la = norm3(e->ea), lb = norm3(e->eb);
	require(c, (la > 0) && (lb > 0) && (e->thick > 0) && (e->epsr >= 1) && (e->ndiva >= 1) && (e->ndivb >= 1) && (e->ndivt >= 1));
	if (c->bad) return KW_BAD;
	if (fabs(inner3(e->ea, e->eb)) > 1e-9 * la * lb) {
		printf("%s\n", "*** dielectric : the two edge vectors must be perpendicular");
		return KW_REPORTED;
	}
	if (e->epsr > 1) p->ndiel++;
	return KW_OK;
}

/* a quad's corners must be coplanar and convex in cyclic order (input_data.c:358-398) */
static int quad_shape(const panel_t *e)
{
	double nrm[3] = {0, 0, 0}, x[3];
	for (int k = 0; k < 4; k++) {
		cross3(&e->v[3 * k], &e->v[3 * ((k + 1) % 4)], x);
		for (int i = 0; i < 3; i++) nrm[i] += x[i];
	}
	const double nl = norm3(nrm);
	if (nl <= 0) return KW_BAD;
	double size = 0;
	for (int k = 0; k < 4; k++) {
		const double *a = &e->v[3 * k], *b = &e->v[3 * ((k + 1) % 4)], *cc = &e->v[3 * ((k + 2) % 4)];
		double u[3], w[3];
		for (int i = 0; i < 3; i++) {
			u[i] = b[i] - a[i];
			w[i] = cc[i] - b[i];
		}
		if (norm3(u) > size) size = norm3(u);
		double off = 0;
		for (int i = 0; i < 3; i++) off += (b[i] - e->v[i]) * nrm[i] / nl;
		if (fabs(off) > 1e-9 * (size + fabs(off))) {
			printf("%s\n", "*** quad : vertices must be coplanar");
			return KW_REPORTED;
		}
		cross3(u, w, x);
		if (inner3(x, nrm) <= 0) {
			printf("%s\n", "*** quad : must be convex (vertices in cyclic order)");
			return KW_REPORTED;
		}
	}
	return KW_OK;
}

/* quad = 4 corners thick sigma na nb */
static int kw_quad(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)arg;
	panel_t *e = (panel_t *)append((void **)&p->panel, p->npanel, &d->panel, sizeof(panel_t));
	if (e == NULL) return KW_BAD;
	e->kind = PANEL_QUAD;
	vec_field(c, e->v, 12);
	e->thick = real_field(c);
	e->sigma = real_field(c);
	e->ndiva = int_field(c);
	e->ndivb = int_field(c);
	require(c, (e->thick > 0) && (e->sigma >= 0) && (e->ndiva >= 1) && (e->ndivb >= 1));
	if (c->bad) return KW_BAD;
	const int shape = quad_shape(e);
	if (shape != KW_OK) return shape;
	p->npanel++;
	return KW_OK;
}

/* disk = centre normal radius thick sigma nrings nsectors.  The reference stops at nsec =
 * POLY_MAX / 2 (its centre cell is a 2 nsec-gon in one cell record); here larger disks, up to
 * DISK_NSEC_MAX sectors, take the centre ring as several sector cells (geometry.c) */
#define DISK_NSEC_MAX 4096
static int kw_disk(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)arg;
	panel_t *e = (panel_t *)append((void **)&p->panel, p->npanel, &d->panel, sizeof(panel_t));
	if (e == NULL) return KW_BAD;
	e->kind = PANEL_DISK;
	vec_field(c, e->org, 3);
	vec_field(c, e->nrm, 3);
	e->radius = real_field(c);
	e->thick = real_field(c);
	e->sigma = real_field(c);
	e->ndiva = int_field(c);
	e->ndivb = int_field(c);
	require(c, (norm3(e->nrm) > 0) && (e->radius > 0) && (e->thick > 0) && (e->sigma >= 0) && (e->ndiva >= 1) &&
	               (e->ndivb >= 3) && (e->ndivb <= DISK_NSEC_MAX));
	if (c->bad) return KW_BAD;
	p->npanel++;
	return KW_OK;
}

static int kw_port(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)arg;
	port_t *e = (port_t *)append((void **)&p->port, p->nport, &d->port, sizeof(port_t));
	if (e == NULL) return KW_BAD;
	e->n1 = node_field(c, p);
	e->n2 = node_field(c, p);
	e->z0 = real_field(c);
	require(c, e->n1 != e->n2);
	if (c->bad) return KW_BAD;
	p->nport++;
	return KW_OK;
}

/* frequency = f0 f1 ndiv [log]: ndiv + 1 points */
static int kw_frequency(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)d;
	(void)arg;
	p->f0 = real_field(c);
	p->f1 = real_field(c);
	const int ndiv = int_field(c);
	if (c->bad) return KW_BAD;
	p->nfreq = ndiv + 1;
	if (more(c, 1) && !strcmp(c->f[c->next], "log")) p->flog = 1;
	require(c, (p->f0 > 0) && (p->f1 >= p->f0) && (ndiv >= 0));
	return c->bad ? KW_BAD : KW_OK;
}

/* integer switches stored at a member offset */
static int kw_flag(peec_t *p, deck_t *d, cursor_t *c, size_t off)
{
	(void)d;
	*(int *)((char *)p + off) = int_field(c);
	return KW_OK;
}

/* real scalars with a lower bound: nodetol > 0, gmin >= 0 (the offset picks the member) */
static int kw_bounded(peec_t *p, deck_t *d, cursor_t *c, size_t off)
{
	(void)d;
	double *v = (double *)((char *)p + off);
	*v = real_field(c);
	require(c, (off == offsetof(peec_t, gmin)) ? (*v >= 0) : (*v > 0));
	return c->bad ? KW_BAD : KW_OK;
}

static int kw_skin(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)d;
	(void)arg;
	p->skin = int_field(c);
	if (p->skin) {
		printf("%s\n", "*** skineffect = 1 is not available in this build (DC resistance only)");
		return KW_REPORTED;
	}
	return KW_OK;
}

/* transient = on [attenuation dB > 0] */
static int kw_transient(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)d;
	(void)arg;
	p->tran = int_field(c);
	if (more(c, 1)) p->tranatt = real_field(c);
	require(c, p->tranatt > 0);
	return c->bad ? KW_BAD : KW_OK;
}

/* planewave = theta phi pol(1 theta, 2 phi) amplitude [phase] */
static int kw_planewave(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)d;
	(void)arg;
	p->pwth = real_field(c);
	p->pwph = real_field(c);
	p->pwpol = int_field(c);
	p->pwamp = real_field(c);
	p->pwphase = more(c, 1) ? real_field(c) : 0;
	require(c, ((p->pwpol == 1) || (p->pwpol == 2)) && (p->pwamp > 0));
	if (c->bad) return KW_BAD;
	p->pw = 1;
	return KW_OK;
}

static int kw_farfield(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)d;
	(void)arg;
	p->ffnth = int_field(c);
	p->ffnph = int_field(c);
	require(c, (p->ffnth >= 2) && (p->ffnph >= 2));
	return c->bad ? KW_BAD : KW_OK;
}

/* compression = on [ACA tolerance > 0] */
static int kw_compression(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)d;
	(void)arg;
	p->compress = int_field(c);
	if (more(c, 1)) p->ctol = real_field(c);
	require(c, p->ctol > 0);
	return c->bad ? KW_BAD : KW_OK;
}

static int kw_groundplane(peec_t *p, deck_t *d, cursor_t *c, size_t arg)
{
	(void)d;
	(void)arg;
	p->gp = 1;
	p->gpz = real_field(c);
	return KW_OK;
}

static const struct {
	const char *key;
	kw_fn fn;
	size_t arg;
} KEYWORDS[] = {
	{"title", kw_title, 0},
	{"resistor", kw_lumped, 0},
	{"capacitor", kw_lumped, 1},
	{"inductor", kw_inductor, 0},
	{"mutual", kw_mutual, 0},
	{"vsource", kw_source, 1},
	{"isource", kw_source, 0},
	{"node", kw_node, 0},
	{"wire", kw_wire, SHAPE_ROUND},
	{"bar", kw_wire, SHAPE_BAR},
	{"plate", kw_plate, 0},
	{"dielectric", kw_dielectric, 0},
	{"quad", kw_quad, 0},
	{"disk", kw_disk, 0},
	{"port", kw_port, 0},
	{"frequency", kw_frequency, 0},
	{"nodetol", kw_bounded, offsetof(peec_t, nodetol)},
	{"gmin", kw_bounded, offsetof(peec_t, gmin)},
	{"skineffect", kw_skin, 0},
	{"capacitance", kw_flag, offsetof(peec_t, capacitance)},
	{"distribution", kw_flag, offsetof(peec_t, dist)},
	{"retardation", kw_flag, offsetof(peec_t, retardation)},
	{"acceleration", kw_flag, offsetof(peec_t, accel)},
	{"grading", kw_flag, offsetof(peec_t, grading)},
	{"transient", kw_transient, 0},
	{"planewave", kw_planewave, 0},
	{"farfield", kw_farfield, 0},
	{"compression", kw_compression, 0},
	{"groundplane", kw_groundplane, 0},
};

/* --------------------------------------------------------------------- the reader -------- */

/* whole-deck rules checked after "end" (input_data.c:540-602) */
static int deck_complete(peec_t *p)
{
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
	const char *why = NULL;
	if (p->nport <= 0) why = "*** no port data";
	else if (p->nfreq <= 0) why = "*** no frequency data";
	else if ((p->nres + p->ncap + p->nind + p->nwire + p->nplate + p->npanel) <= 0)
		why = "*** no element data (resistor/capacitor/inductor/wire/bar/plate/polygon/disk)";
	else if ((p->ndiel > 0) && !p->capacitance) why = "*** dielectric requires capacitance = 1";
	else if (p->compress && p->accel) why = "*** compression = 1 and acceleration = 1 are mutually exclusive";
	else if (p->compress && ((p->nwire + p->nplate + p->npanel) <= 0))
		why = "*** compression = 1 requires conductor geometry (wire/bar/plate/quad/disk)";
	if (why) {
		printf("%s\n", why);
		return 1;
	}
	return 0;
}

/* the next statement line: blank lines (at most one character with its newline) and lines
 * starting with '#' are passed over; the line end goes (two characters when the line holds
 * "\r\n", else one when it holds '\r' or '\n').  0 at end of file. */
static int read_line(FILE *fp, char *buf, size_t cap)
{
	for (;;) {
		if (fgets(buf, (int)cap, fp) == NULL) return 0;
		const size_t n = strlen(buf);
		if ((n <= 1) || (buf[0] == '#')) continue;
		if (strstr(buf, "\r\n") != NULL) buf[n - 2] = '\0';
		else if (strpbrk(buf, "\r\n") != NULL) buf[n - 1] = '\0';
		return 1;
	}
}

int input_data(FILE *fp, peec_t *p)
{
	char raw[BUFSIZ], line[BUFSIZ];
	char *tok[MAX_FIELDS + 1];
	deck_t d;
	memset(&d, 0, sizeof(d));
	memset(p, 0, sizeof(peec_t));
	p->nodetol = 1e-8;
	p->refnode = -1;
	p->tranatt = 40;
	p->ctol = 1e-6;

	int header = 1;   /* the first significant line must be the format header */
	while (read_line(fp, raw, sizeof(raw))) {
		if (!strncmp(raw, "end", 3)) break;
		snprintf(line, sizeof(line), "%s", raw);
		const int nt = tokenize(line, " \t", tok, MAX_FIELDS);
		if (header) {
			if ((nt == 0) || strcmp(tok[0], "OpenPEEC")) {
				printf("%s\n", "*** not OpenPEEC data");
				return 1;
			}
			header = 0;
			continue;
		}
		if ((nt < 3) || strcmp(tok[1], "=")) continue;   /* not "key = value ..." */
		for (size_t k = 0; k < sizeof(KEYWORDS) / sizeof(KEYWORDS[0]); k++) {
			if (strcmp(tok[0], KEYWORDS[k].key)) continue;
			cursor_t c = {tok + 2, nt - 2, 0, 0};
			d.line = raw;
			const int rc = KEYWORDS[k].fn(p, &d, &c, KEYWORDS[k].arg);
			if (rc == KW_BAD) printf("*** invalid %s data\n", KEYWORDS[k].key);
			if (rc != KW_OK) return 1;
			break;
		}
		/* keys not in the table are ignored (forward compatibility) */
	}
	return deck_complete(p);
}

void peec_free(peec_t *p)
{
	if (p == NULL) return;
	peec_gpu_release(p);
	void *owned[] = {p->plate, p->panel, p->diel, p->res, p->cap, p->ind, p->mut, p->src, p->port, p->wire, p->ncid,
	                 p->ncxyz, p->seg, p->chg, p->chgnode, p->gid, p->gxyz, p->lp, p->cellid, p->cellof, p->carea,
	                 p->csoff, p->csidx, p->cpt, p->cmat, p->nodemap, p->zin, p->zmat, p->smat, p->segi, p->cellq,
	                 p->voc, p->segipw};
	for (size_t k = 0; k < sizeof(owned) / sizeof(owned[0]); k++) free(owned[k]);
	memset(p, 0, sizeof(peec_t));
}
