/* Place Fortran COMMON blocks (about 50 GB of zero-initialised, never
   touched capacity in the reference's compile-time array sizes) after
   .bss so that the runtime's 32-bit PC-relative relocations still reach
   its own .bss.  Pure layout: no symbols are defined or replaced here. */
SECTIONS { .lbss (NOLOAD) : { *(COMMON) *(.lbss .lbss.*) } } INSERT AFTER .bss;
