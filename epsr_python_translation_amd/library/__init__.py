"""Drop-in replacements, by module and name, for the reference's library routines on the
hot path (src/epsr26/library/): ato_manipulation_rev4.atogen/stog/fmolec1,
molfit.molfit/en_change/en_mole_pair, auxiliary_routines.ran1/smoop."""
