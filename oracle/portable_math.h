/* portable elementary functions (placeholder) */
