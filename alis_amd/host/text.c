/*
 * Text outputs: writeTxt, writeRow (src/output.c:104-174) and the column spectra of phaser
 * arrays, writePoyntingSpectrum and writeFarFieldSpectrum (src/phaser.c:248-295).
 *
 * An output is keyed by its world and its formatted name.  Its file, W->ID + name + ".txt",
 * is created (truncated) by the first write and stays open; it is flushed when five minutes
 * have passed since its last flush, when its world is released, and by the C library at
 * exit.  A name that starts with '/' first creates the directory W->ID; a name that ends
 * with '/' creates the directory W->ID + name.
 *
 * writeTxt's arguments are one list: the name's format consumes its own arguments and the
 * next argument is the text's format, which consumes the rest.  The text is written as
 * formatted (the reference passes it through printf a second time, so a '%' that the first
 * formatting produced would be read as a conversion; here it is written as a '%').
 *
 * How far vsnprintf leaves a va_list it was given is not defined by C (glibc's advances it
 * past the arguments it formatted; a vsnprintf interposed by another library loaded into the
 * process may not), so the functions here that read arguments after a name's format format
 * the name from a copy and step over its arguments themselves (alisFormatName).
 */
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include "internal.h"

typedef struct {
	world W;
	char *name;
	FILE *file;
	double flushed;      /* seconds on the monotonic clock */
} textOut;

/* Format a name into out from a copy of *ap, then advance *ap past the arguments the format
 * names: glibc conversions with flags, widths and precisions (also '*'), length modifiers
 * hh h l ll q L j z Z t; positional arguments ("%1$d") are not supported. */
void alisFormatName(char *out, size_t size, const char *format, va_list *ap)
{
	va_list copy;
	va_copy(copy, *ap);
	vsnprintf(out, size, format, copy);
	va_end(copy);
	for (const char *c = format; *c; c++) {
		if (*c != '%') continue;
		if (*++c == '%') continue;
		while (*c && strchr("-+ #0'I", *c)) c++;
		if (*c == '*') (void)va_arg(*ap, int), c++;
		while (*c >= '0' && *c <= '9') c++;
		if (*c == '.') {
			c++;
			if (*c == '*') (void)va_arg(*ap, int), c++;
			while (*c >= '0' && *c <= '9') c++;
		}
		int wide = 0, longDouble = 0;
		for (; *c && strchr("hlqLjzZt", *c); c++) {
			if (*c == 'L') longDouble = 1;
			else if (*c != 'h') wide = 1;
		}
		if (!*c) break;
		if (strchr("diouxXc", *c)) {
			if (wide) (void)va_arg(*ap, long);
			else (void)va_arg(*ap, int);
		}
		else if (strchr("eEfFgGaA", *c)) {
			if (longDouble) (void)va_arg(*ap, long double);
			else (void)va_arg(*ap, double);
		}
		else if (strchr("sSpn", *c)) (void)va_arg(*ap, void *);
	}
}

static textOut *outs;
static int nOuts, capOuts;

static double seconds(void)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return t.tv_sec + 1e-9 * t.tv_nsec;
}

static textOut *outputOf(world W, const char *name)
{
	for (int t = 0; t < nOuts; t++)
		if (outs[t].W == W && !strcmp(outs[t].name, name)) return outs + t;
	if (nOuts == capOuts) {
		capOuts = capOuts ? 2 * capOuts : 16;
		outs = realloc(outs, capOuts * sizeof *outs);
	}
	char path[2048];
	int n = snprintf(path, sizeof path, "%s%s", W->ID, name);
	if (name[0] == '/') mkdir(W->ID, 0755);
	if (n > 0 && path[n - 1] == '/') mkdir(path, 0755);
	snprintf(path + n, sizeof path - n, ".txt");
	textOut *t = outs + nOuts++;
	*t = (textOut){W, strdup(name), fopen(path, "w"), seconds()};
	return t;
}

static void put(world W, const char *name, const char *text)
{
	textOut *t = outputOf(W, name);
	if (!t->file) return;
	fputs(text, t->file);
	double now = seconds();
	if (now - t->flushed >= 300) {
		fflush(t->file);
		t->flushed = now;
	}
}

/* Close the outputs of a world that is being released. */
void alisForgetTexts(world W)
{
	int kept = 0;
	for (int t = 0; t < nOuts; t++) {
		if (outs[t].W != W) {
			outs[kept++] = outs[t];
			continue;
		}
		if (outs[t].file) fclose(outs[t].file);
		free(outs[t].name);
	}
	nOuts = kept;
}

void (writeTxt)(world W, char *format, ...)
{
	char name[1024], small[1024];
	va_list ap, again;
	va_start(ap, format);
	alisFormatName(name, sizeof name, format, &ap);
	const char *body = va_arg(ap, const char *);
	va_copy(again, ap);
	int n = vsnprintf(small, sizeof small, body, ap);
	char *text = small;
	if (n >= (int)sizeof small && (text = malloc(n + 1))) vsnprintf(text, n + 1, body, again);
	va_end(again);
	va_end(ap);
	if (n >= 0 && text) put(W, name, text);
	if (text != small) free(text);
}

/* A growable line of tab-separated %g numbers. */
typedef struct {
	char *s;
	size_t len, cap;
} line;

static void append(line *l, double v)
{
	char num[40];
	int n = snprintf(num, sizeof num, l->len ? "\t%g" : "%g", v);
	if (l->len + n + 2 > l->cap) {
		l->cap = 2 * (l->len + n + 2);
		l->s = realloc(l->s, l->cap);
	}
	memcpy(l->s + l->len, num, n + 1);
	l->len += n;
}

static void putLine(world W, const char *name, line *l)
{
	l->s[l->len++] = '\n';
	l->s[l->len] = 0;
	put(W, name, l->s);
	l->len = 0;
}

/* One row: the numbers up to the INT2NAN(1) sentinel; no numbers, no row (and no file). */
void (writeRow)(world W, char *format, ...)
{
	char name[1024];
	line l = {0};
	va_list ap;
	va_start(ap, format);
	alisFormatName(name, sizeof name, format, &ap);
	for (double v = va_arg(ap, double); !NAN2INT(v); v = va_arg(ap, double)) append(&l, v);
	va_end(ap);
	if (l.len) putLine(W, name, &l);
	free(l.s);
}

/* One row per frequency of the first phaser array (to its zero terminator): the wavelength
 * 1/f, then value(array m's phaser of that frequency) for every array of the NULL-terminated
 * list that follows the name's arguments. */
static void columns(world W, const char *format, va_list *ap, float value(phaser, const void *), const void *arg)
{
	char name[1024];
	alisFormatName(name, sizeof name, format, ap);
	phaser *P = 0;
	int M = 0;
	for (phaser p; (p = va_arg(*ap, phaser)); M++) {
		if (!(M & (M - 1))) P = realloc(P, (M ? 2 * M : 1) * sizeof *P);
		P[M] = p;
	}
	line l = {0};
	for (int n = 0; M && P[0][n].f; n++) {
		append(&l, 1 / P[0][n].f);
		for (int m = 0; m < M; m++) append(&l, value(P[m] + n, arg));
		putLine(W, name, &l);
	}
	free(l.s);
	free(P);
}

static float poyntingOf(phaser P, const void *unused)
{
	(void)unused;
	return phaserPoynting(P);
}

typedef struct {
	tfield (*pol)(phaser, float, float);
	float theta, phi;
} direction;

static float intensityOf(phaser P, const void *arg)
{
	const direction *d = arg;
	return farFieldI(P, d->pol, d->theta, d->phi);
}

void (writePoyntingSpectrum)(world W, char *format, ...)
{
	va_list ap;
	va_start(ap, format);
	columns(W, format, &ap, poyntingOf, 0);
	va_end(ap);
}

void (writeFarFieldSpectrum)(world W, tfield pol(phaser, float, float), float theta, float phi, char *format, ...)
{
	direction d = {pol, theta, phi};
	va_list ap;
	va_start(ap, format);
	columns(W, format, &ap, intensityOf, &d);
	va_end(ap);
}
