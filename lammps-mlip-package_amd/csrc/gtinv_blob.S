/* Embeds the generalized Clebsch-Gordan table (data/gtinv_table.txt) in the
   library so that the engine needs no data file at run time.  The build
   passes the data directory to the assembler include path. */
    .section .rodata
    .balign 16
    .global mlip_gtinv_table_begin
    .global mlip_gtinv_table_end
mlip_gtinv_table_begin:
    .incbin "gtinv_table.txt"
mlip_gtinv_table_end:
    .byte 0
    .section .note.GNU-stack,"",@progbits
