"""Drop-in replacements for the reference's epsr_standalone_rev3 routines on the hot path
(src/epsr26/programs/epsr_standalone_rev3.py): gensq, expand, fit_expansion,
calc_pwrexpq, generate_rp, generate_ep, ptgnsq, and the inita/initpt table setup."""
