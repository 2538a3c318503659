/* The crossing predictor's generated tables (predictor_tables.bin, written
 * by extract_predictor_tables.py), embedded read-only in the library. */
        .section .rodata
        .balign 16
        .globl jpxlg_predictor_tables
        .type jpxlg_predictor_tables, @object
jpxlg_predictor_tables:
        .incbin "predictor_tables.bin"
        .globl jpxlg_predictor_tables_end
jpxlg_predictor_tables_end:
        .size jpxlg_predictor_tables, jpxlg_predictor_tables_end - jpxlg_predictor_tables
        .section .note.GNU-stack,"",@progbits
